// The schedule of the AV sweeps of a 3D step with a periodic z axis run as a
// pipeline over z slabs (host code, no HIP): the one-walk x-y AV sweeps in
// slabs on one stream and the z walk in slabs on another.  capi.cpp
// launches what the plan says; tests/cpp/test_tail_plan.cpp checks plans
// against the kernels' plane accesses.
//
// Planes are the array's z indices 0 .. Kt-1 with margin mk; nk = Kt - 2 mk
// interior planes; a ghost plane of a periodic axis is read at its source
// (q < mk: q + nk; q >= Kt - mk: q - nk).  With a fixed z boundary the swept
// fields' ghost planes are filled (from the fixed values) before the
// pipeline starts, and every plane is read as itself.
//
//   x-y slab s (planes xy0[s] .. xy0[s] + xyn[s] - 1; a plane index past
//     the top z ring plane Kt - mk continues at the bottom ring plane
//     mk - 1): reads the post-RK state at each plane's source, writes the
//     swept fields at its interior planes and cc at all of them.
//   z slab s (the z walk's chunks [s ? c1[s-1] : 0, c1[s]) of zch planes):
//     reads the swept fields at its planes +-2 (sources), cc at its planes
//     +-1 (no mapping: the x-y walk writes the ring planes), writes qq at
//     its planes.
//
// x-y slab s covers the planes z slab s reads that no earlier x-y slab
// covered; with a periodic z axis slab 0 also the two top interior planes
// and the top ring plane, which z slab 0 reads as its periodic ghosts (so
// on a short grid the last slabs can be empty: nothing is left for them).
// Each z slab waits for the latest x-y slab it conflicts with; launches of
// one stream run in order.
#pragma once

#include <algorithm>
#include <vector>

namespace misohip {

constexpr int TAIL_SLABS = 16;

struct TailPlan {
  int ns = 0;                      // slabs
  int c1[TAIL_SLABS];              // z slab s: chunks [s ? c1[s-1] : 0, c1[s])
  int xy0[TAIL_SLABS], xyn[TAIL_SLABS];
  int z_xy[TAIL_SLABS];            // z slab s waits for x-y slab z_xy[s] (-1: none)
};

// false: no pipeline for this shape (fewer than two slabs, or a grid of
// fewer than 8 planes)
inline bool tail_plan(int Kt, int mk, int zch, int znch, bool periodic_z, TailPlan *pl) {
  const int nk = Kt - 2 * mk;
  const int ns = std::min(znch, TAIL_SLABS);
  if (nk < 8 || zch < 1 || ns < 2) return false;
  TailPlan &P = *pl;
  P.ns = ns;
  for (int s = 0; s < ns; ++s) P.c1[s] = (int)((long)(s + 1) * znch / ns);
  // periodic: the first of the two top interior planes, which slab 0 takes;
  // fixed: past the top ring plane
  const int top0 = periodic_z ? Kt - mk - 2 : Kt - mk + 1;
  auto src = [&](int q) {
    return !periodic_z ? q : q < mk ? q + nk : q >= Kt - mk ? q - nk : q;
  };
  using Mask = std::vector<char>;
  auto hit = [&](const Mask &x, const Mask &y) {
    for (int q = 0; q < Kt; ++q)
      if (x[q] && y[q]) return true;
    return false;
  };
  std::vector<Mask> xy_wd(ns, Mask(Kt, 0)), xy_wc(ns, Mask(Kt, 0));
  for (int s = 0; s < ns; ++s) {
    const int hi = s + 1 < ns ? std::min(mk + P.c1[s] * zch + 2, top0) : top0;
    const int lo = s == 0 ? mk - 1 : std::min(mk + P.c1[s - 1] * zch + 2, top0);
    P.xy0[s] = s == 0 && periodic_z ? top0 : lo;
    P.xyn[s] = s == 0 && periodic_z ? 3 + hi - lo : std::max(0, hi - lo);
    for (int t = 0; t < P.xyn[s]; ++t) {
      int q = P.xy0[s] + t;
      if (q > Kt - mk) q -= nk + 2;
      if (q >= mk && q < Kt - mk) xy_wd[s][q] = 1;
      xy_wc[s][q] = 1;
    }
  }
  for (int s = 0; s < ns; ++s) {
    Mask rd(Kt, 0), rc(Kt, 0);
    const int p0 = mk + (s == 0 ? 0 : P.c1[s - 1]) * zch, p1 = mk + std::min(P.c1[s] * zch, nk);
    for (int q = p0 - 2; q < p1 + 2; ++q) rd[src(q)] = 1;
    for (int q = p0 - 1; q <= p1; ++q) rc[q] = 1;
    P.z_xy[s] = -1;
    for (int t = 0; t < ns; ++t)
      if (hit(xy_wd[t], rd) || hit(xy_wc[t], rc)) P.z_xy[s] = t;
  }
  return true;
}

}  // namespace misohip
