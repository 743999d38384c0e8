// CPU check of the AV slab pipeline's schedule (miso_amd/csrc/tail_plan.hpp):
// for many grid heights and chunkings, with a periodic and with a fixed z
// axis, every pair of launches on the two
// streams that touch the same plane of the same buffer, one of them
// writing, is ordered by the plan's waits and the streams' own order, x-y
// slab before z slab (the order of the sequential walks); and the slabs
// together write every plane exactly once.  The launches' plane
// accesses are stated here from the kernels (av_sweep.hip), not taken from
// the plan.  Built by tests/cpp/Makefile with g++; run by
// tests/test_tail_plan.py.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../miso_amd/csrc/tail_plan.hpp"

namespace {

enum Buf { SWEPT = 0, RES = 1, QQ = 2, CC = 3, NBUF = 4 };  // RES: the post-RK state

struct Launch {
  int pass;  // 1 x-y, 2 z: the sequential order
  std::vector<std::vector<char>> rd, wr;  // [buffer][plane]
};

int check(int Kt, int mk, int zch, int znch, bool pz, char *msg, int n) {
  using misohip::TailPlan;
  TailPlan P;
  if (!misohip::tail_plan(Kt, mk, zch, znch, pz, &P)) return 0;  // no pipeline: nothing to check
  const int nk = Kt - 2 * mk;
  // a ghost plane: its periodic source; with a fixed z boundary itself (its
  // swept fields filled before the pipeline, by no launch of it)
  auto src = [&](int q) { return !pz ? q : q < mk ? q + nk : q >= Kt - mk ? q - nk : q; };
  auto blank = [&]() {
    Launch L;
    L.rd.assign(NBUF, std::vector<char>(Kt, 0));
    L.wr.assign(NBUF, std::vector<char>(Kt, 0));
    return L;
  };
  std::vector<Launch> launches;
  std::vector<int> id_xy(P.ns), id_z(P.ns);
  // x-y slab s: each plane it walks reads the post-RK state in that plane (a
  // z ring plane: its source plane, for cc only) and writes the swept fields
  // on interior planes, cc on every plane
  for (int s = 0; s < P.ns; ++s) {
    Launch L = blank();
    L.pass = 1;
    for (int t = 0; t < P.xyn[s]; ++t) {
      int q = P.xy0[s] + t;
      if (q > Kt - mk) q -= nk + 2;
      if (q < mk - 1 || q > Kt - mk) return std::snprintf(msg, n, "x-y plane %d off the walk", q), 1;
      L.rd[RES][src(q)] = 1;
      if (q >= mk && q < Kt - mk) L.wr[SWEPT][q] = 1;
      L.wr[CC][q] = 1;
    }
    id_xy[s] = (int)launches.size();
    launches.push_back(L);
  }
  // z slab s: walks chunks [z0, z1) of zch planes; the flux at a face reads
  // the swept fields two planes either side of the cells (sources) and cc
  // one plane either side (the ring planes themselves); writes qq
  for (int s = 0; s < P.ns; ++s) {
    Launch L = blank();
    L.pass = 2;
    const int z0 = s == 0 ? 0 : P.c1[s - 1];
    for (int q = mk + z0 * zch; q < mk + P.c1[s] * zch && q < Kt - mk; ++q) {
      for (int d = -2; d <= 2; ++d) L.rd[SWEPT][src(q + d)] = 1;
      for (int d = -1; d <= 1; ++d) L.rd[CC][q + d] = 1;
      L.wr[QQ][q] = 1;
    }
    id_z[s] = (int)launches.size();
    launches.push_back(L);
  }
  const int N = (int)launches.size();
  // happens-before: stream order and the plan's waits, closed transitively
  std::vector<std::vector<char>> hb(N, std::vector<char>(N, 0));
  for (int s = 1; s < P.ns; ++s) hb[id_xy[s - 1]][id_xy[s]] = 1;
  for (int s = 1; s < P.ns; ++s) hb[id_z[s - 1]][id_z[s]] = 1;
  for (int s = 0; s < P.ns; ++s)
    if (P.z_xy[s] >= 0) hb[id_xy[P.z_xy[s]]][id_z[s]] = 1;
  for (int k = 0; k < N; ++k)
    for (int i = 0; i < N; ++i)
      if (hb[i][k])
        for (int j = 0; j < N; ++j)
          if (hb[k][j]) hb[i][j] = 1;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      if (launches[i].pass >= launches[j].pass) continue;  // i's pass comes first
      for (int b = 0; b < NBUF; ++b)
        for (int q = 0; q < Kt; ++q) {
          const bool conflict = (launches[i].wr[b][q] && (launches[j].rd[b][q] || launches[j].wr[b][q])) ||
                                (launches[i].rd[b][q] && launches[j].wr[b][q]);
          if (conflict && !hb[i][j])
            return std::snprintf(msg, n,
                                 "Kt %d zch %d znch %d periodic %d: launch %d (pass %d) and %d "
                                 "(pass %d) share buffer %d plane %d unordered",
                                 Kt, zch, znch, (int)pz, i, launches[i].pass, j,
                                 launches[j].pass, b, q),
                   1;
        }
    }
  // coverage: every plane written once per pass and buffer
  const int want[NBUF][3] = {{0, 1, 0}, {0, 0, 0}, {0, 0, 1}, {0, 1, 0}};
  for (int b = 0; b < NBUF; ++b)
    for (int pass = 1; pass < 3; ++pass) {
      if (!want[b][pass]) continue;
      for (int q = 0; q < Kt; ++q) {
        int w = 0;
        for (const Launch &L : launches)
          if (L.pass == pass) w += L.wr[b][q];
        const bool interior = q >= mk && q < Kt - mk, ring = q == mk - 1 || q == Kt - mk;
        const int expect = interior || (b == CC && ring) ? 1 : 0;
        if (w != expect)
          return std::snprintf(msg, n,
                               "Kt %d zch %d znch %d periodic %d: buffer %d plane %d written %d "
                               "times",
                               Kt, zch, znch, (int)pz, b, q, w),
                 1;
      }
    }
  return 0;
}

}  // namespace

// Every plan over a sweep of heights and chunkings; 0 when all hold, else 1
// with the first failure in msg.  *plans: how many shapes had a pipeline.
extern "C" int tail_plan_check(char *msg, int n, int *plans) {
  const int mk = 2;
  *plans = 0;
  for (int nk = 8; nk <= 1100; nk += (nk < 100 ? 1 : 29)) {
    for (int zreq : {4, 5, 7, 8, 16, 32, 64, 128, 256})
      for (bool pz : {true, false}) {
        const int znch = (nk + zreq - 1) / zreq, zch = (nk + znch - 1) / znch;
        misohip::TailPlan P;
        if (misohip::tail_plan(nk + 2 * mk, mk, zch, znch, pz, &P)) ++*plans;
        if (check(nk + 2 * mk, mk, zch, znch, pz, msg, n)) return 1;
      }
  }
  msg[0] = 0;
  return 0;
}
