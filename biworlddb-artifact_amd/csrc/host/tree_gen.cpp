// tree_gen.cpp -- see tree_gen.hpp.
#include "tree_gen.hpp"

#include <cstdio>

namespace bwg {
namespace {

constexpr uint32_t kTieCap = 1u << 16;

bool fail(char* err, size_t errlen, const char* what, int rc) {
  if (err && errlen) std::snprintf(err, errlen, "%s: %s", what, hipGetErrorString(hipError_t(rc)));
  return false;
}

// Elapsed device time between two events recorded on st (synchronises on the second).
double elapsed_ms(hipEvent_t a, hipEvent_t b) {
  float t = 0;
  if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess) return -1;
  return double(t);
}

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  Events() {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

}  // namespace

bool device_zipf(uint64_t m, double s, DeviceZipf& z, hipStream_t st, char* err, size_t errlen) {
  z.arena.release();
  z.arena.err = hipSuccess;
  z.m = m;
  z.cum = z.arena.upload(zipf_cumulative(m, s));
  z.guide = z.arena.scratch<uint32_t>(BWT_ZIPF_GUIDE_ENTRIES);
  if (!z.arena.ok()) return fail(err, errlen, "zipf table", int(z.arena.err));
  const int rc = bwt_gen_guide_launch(z.cum, m, z.guide, st);
  return rc ? fail(err, errlen, "zipf guide", rc) : true;
}

bool device_generate_tree(const TreeSpec& spec, const std::vector<uint32_t>& dkey, DeviceTree& out,
                          hipStream_t st, char* err, size_t errlen) {
  out.arena.release();
  out.arena.err = hipSuccess;
  DeviceZipf z;
  BwtGenSpec g{};
  g.cfg = spec.cfg;
  g.U = spec.U;
  g.Q = spec.Q;
  g.S = spec.S;
  g.N = spec.N;
  g.seed = spec.seed;
  if (spec.cfg != 1) {
    const uint64_t m = spec.cfg == 3 ? uint64_t(dkey.size()) : spec.N;
    if (!device_zipf(m, spec.cfg == 3 ? 0.9 : 1.1, z, st, err, errlen)) return false;
    g.zipf_cum = z.cum;
    g.zipf_n = m;
    g.zipf_guide = z.guide;
    if (spec.cfg == 3) g.dkey = out.arena.upload(dkey);
  }
  const size_t nu = size_t(spec.S) * size_t(spec.U), nq = size_t(spec.S) * size_t(spec.Q);
  out.key = out.arena.scratch<uint32_t>(nu);
  out.delta = out.arena.scratch<int64_t>(nu);
  out.qlo = out.arena.scratch<uint32_t>(nq);
  out.qhi = out.arena.scratch<uint32_t>(nq);
  BwtGenTies t{out.arena.scratch<uint32_t>(kTieCap), out.arena.scratch<uint32_t>(1), kTieCap};
  if (!out.arena.ok()) return fail(err, errlen, "tree arrays", int(out.arena.err));
  Events ev;
  int rc = int(hipEventRecord(ev.a, st));
  if (!rc) rc = bwt_gen_updates_launch(&g, out.key, out.delta, st);
  if (!rc) rc = bwt_gen_state_queries_launch(&g, out.qlo, out.qhi, t, st);
  if (!rc) rc = int(hipEventRecord(ev.b, st));
  if (rc) return fail(err, errlen, "tree generation", rc);
  out.gpu_ms = elapsed_ms(ev.a, ev.b);
  if (hipMemcpy(&out.ties, t.count, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) out.ties = 0;
  rc = bwt_gen_settle(&g, 0, out.qlo, out.qhi, t, st);
  return rc ? fail(err, errlen, "tree ties", rc) : true;
}

bool device_generate_stream(const TreeSpec& tree, const RequestStream& scan, DeviceStream& out, hipStream_t st,
                            char* err, size_t errlen,
                            const std::vector<uint32_t>& dkey) {
  out.arena.release();
  out.arena.err = hipSuccess;
  const StreamSpec& ss = scan.spec;
  const int64_t R = ss.nreq;
  if (int64_t(scan.req_q0.size()) != R + 1 || int64_t(scan.req_app.size()) != R) {
    if (err && errlen) std::snprintf(err, errlen, "stream generation: scan arrays missing");
    return false;
  }
  out.nq = scan.req_q0[size_t(R)];
  out.napp = R ? scan.req_app[size_t(R) - 1] + scan.req_kind[size_t(R) - 1] : 0;
  DeviceZipf z;
  BwtGenSpec g{};
  g.cfg = tree.cfg;
  g.U = tree.U;
  g.Q = tree.Q;
  g.S = tree.S;
  g.N = tree.N;
  g.seed = tree.seed;
  if (ss.form == 5) {
    // over a cfg-3 tree the stream draws distinct keys by the tree's own Zipf(0.9) law
    const bool ranked = tree.cfg == 3;
    if (ranked && uint64_t(dkey.size()) != tree.N) {
      if (err && errlen) std::snprintf(err, errlen, "stream generation: the cfg-3 tree's distinct keys are missing");
      return false;
    }
    if (!device_zipf(tree.N, ranked ? 0.9 : 1.1, z, st, err, errlen)) return false;
    g.zipf_cum = z.cum;
    g.zipf_n = tree.N;
    g.zipf_guide = z.guide;
    if (ranked) g.dkey = out.arena.upload(dkey);
  }
  const int64_t* req_q0 = ss.form == 5 ? out.arena.upload(scan.req_q0) : nullptr;
  const int64_t* req_app = ss.form == 5 ? out.arena.upload(scan.req_app) : nullptr;
  out.qstate = out.arena.scratch<int32_t>(size_t(out.nq));
  out.qlo = out.arena.scratch<uint32_t>(size_t(out.nq));
  out.qhi = out.arena.scratch<uint32_t>(size_t(out.nq));
  const size_t na = size_t(out.napp) * size_t(ss.append_U);
  out.app_parent = out.arena.scratch<int32_t>(size_t(out.napp));
  out.app_key = out.arena.scratch<uint32_t>(na);
  out.app_delta = out.arena.scratch<int64_t>(na);
  BwtGenTies t{out.arena.scratch<uint32_t>(kTieCap), out.arena.scratch<uint32_t>(1), kTieCap};
  if (!out.arena.ok()) return fail(err, errlen, "stream arrays", int(out.arena.err));
  Events ev;
  int rc = int(hipEventRecord(ev.a, st));
  if (!rc)
    rc = bwt_gen_stream_launch(&g, ss.form, req_q0, req_app, R, out.nq, out.qstate, out.qlo, out.qhi, t, st);
  if (!rc && out.napp)
    rc = bwt_gen_appends_launch(&g, out.napp, ss.append_U, out.app_parent, out.app_key, out.app_delta, st);
  if (!rc) rc = int(hipEventRecord(ev.b, st));
  if (rc) return fail(err, errlen, "stream generation", rc);
  out.gpu_ms = elapsed_ms(ev.a, ev.b);
  if (hipMemcpy(&out.ties, t.count, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) out.ties = 0;
  rc = bwt_gen_settle(&g, 1, out.qlo, out.qhi, t, st);
  return rc ? fail(err, errlen, "stream ties", rc) : true;
}

}  // namespace bwg
