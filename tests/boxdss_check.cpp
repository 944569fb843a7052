// CPU check of the structured-box group construction (gf_boxdss.hpp) against
// the GGFX maps of gf_ggfx_build: for every node of every case, the members
// from box_members_own (64- and 32-bit indices) equal box_group and, where
// the box's lattice groups are the mesh's GGFX groups (every case listed with
// expect_ggfx = 1), the GGFX group in the same ascending order; the own slot
// holds the node itself.  At p = 7 the LDS halo of the stage kernels is
// simulated too: each element's own values are copied to a halo array and
// its 64 lanes fill the rest with the global indices their 9 partner loads
// read (Halo8Rim), and for every node the halo
// slots of halo8_order, read in order with the zero slot skipped, must give
// box_group's members (and only the z faces may use the last four terms) on
// every box where box_halo8_ok holds, and fail somewhere on the others.
// Exit status 0 when everything holds.
//
// The two cases with expect_ggfx = 0 mix periodic and non-periodic
// directions: GGridBox::periodize (src/cdg/grid/ggrid_box.ipp:753-802) also
// moves the nodes of a periodic face that lie on the far face of a
// non-periodic direction, so GGFX joins them with the near face and the groups
// are not lattice groups; gf_mesh_create detects this and keeps the CSR DSS.
#define __host__
#define __device__
#define __forceinline__ inline
#define HIP_INCLUDE_HIP_HIP_RUNTIME_H
#include <cstdint>
#include <cstdio>
#include <vector>

#include "gf_boxdss.hpp"
#include "geoflow_hip.h"

int main() {
  struct Case {
    int p, ne[3], bdy[6], expect_ggfx;
  } cases[] = {
      {3, {3, 4, 5}, {6, 6, 6, 6, 6, 6}, 1}, {2, {1, 1, 1}, {6, 6, 6, 6, 6, 6}, 1},
      {7, {3, 3, 3}, {0, 0, 0, 0, 0, 0}, 1}, {1, {2, 2, 2}, {6, 6, 6, 6, 6, 6}, 1},
      {3, {1, 2, 2}, {6, 6, 6, 6, 6, 6}, 1}, {1, {1, 1, 2}, {6, 6, 6, 6, 6, 6}, 1},
      {4, {2, 2, 1}, {6, 6, 6, 6, 6, 6}, 1}, {2, {1, 3, 1}, {0, 0, 0, 0, 0, 0}, 1},
      {6, {2, 1, 2}, {6, 6, 6, 6, 6, 6}, 1}, {1, {5, 1, 1}, {6, 6, 6, 6, 6, 6}, 1},
      {5, {2, 3, 2}, {0, 0, 0, 0, 0, 0}, 1}, {4, {2, 1, 3}, {6, 6, 0, 0, 6, 6}, 0},
      {2, {4, 3, 2}, {6, 6, 0, 0, 0, 0}, 0}, {7, {3, 2, 2}, {6, 6, 6, 6, 6, 6}, 1},
      {7, {1, 1, 1}, {6, 6, 6, 6, 6, 6}, 1}, {7, {2, 1, 3}, {6, 6, 6, 6, 6, 6}, 1},
      {7, {2, 2, 3}, {0, 0, 0, 0, 0, 0}, 1}, {7, {1, 3, 2}, {6, 6, 6, 6, 6, 6}, 1},
      {7, {3, 3, 1}, {6, 6, 6, 6, 6, 6}, 1}, {7, {4, 4, 4}, {6, 6, 6, 6, 6, 6}, 1}};
  int bad = 0;
  for (const Case &c : cases) {
    const int p3[3] = {c.p, c.p, c.p}, N = c.p + 1;
    const double P0[3] = {0, 0, 0}, P1[3] = {1, 1.5, 2};
    const size_t nel = (size_t)c.ne[0] * c.ne[1] * c.ne[2], nn = (size_t)N * N * N, n = nel * nn;
    std::vector<double> xp[3];
    double *xpp[3];
    for (int d = 0; d < 3; d++) {
      xp[d].resize(n);
      xpp[d] = xp[d].data();
    }
    double mnd = 0;
    if (gf_box_grid(3, p3, c.ne, P0, P1, c.bdy, nullptr, xpp, nullptr, nullptr, nullptr, &mnd))
      return 2;
    std::vector<int64_t> offs(n + 1), mem;
    size_t nm = 0;
    const double *cx[3] = {xpp[0], xpp[1], xpp[2]};
    gf_ggfx_build(n, 3, cx, 0.001 * mnd, 8, offs.data(), nullptr, &nm);
    mem.resize(nm);
    gf_ggfx_build(n, 3, cx, 0.001 * mnd, 8, offs.data(), mem.data(), &nm);
    gf::BoxTopo t = {};
    for (int d = 0; d < 3; d++) {
      t.ne[d] = c.ne[d];
      t.per[d] = c.bdy[2 * d] == GF_BDY_PERIODIC && c.bdy[2 * d + 1] == GF_BDY_PERIODIC;
    }
    size_t wrong = 0, ggfx_diff = 0;
    for (size_t q = 0; q < n; q++) {
      const size_t e = q / nn, r = q - e * nn;
      const int i = (int)(r % N), j = (int)((r / N) % N), k = (int)(r / ((size_t)N * N));
      int ec[3];
      gf::box_elem_coords(t, e, ec);
      size_t g[8];
      const int cnt = gf::box_group(t, N, ec, q, i, j, k, g);
      const gf::BoxElemDeltas de = gf::box_elem_deltas(t, N, ec);
      int64_t g64[8];
      int32_t g32[8];
      int own64 = -1, own32 = -1;
      const int c64 = gf::box_members_own<int64_t>(de, N, (int64_t)q, i, j, k, g64, &own64);
      const int c32 = gf::box_members_own<int32_t>(de, N, (int32_t)q, i, j, k, g32, &own32);
      bool ok = c64 == cnt && c32 == cnt && own64 == own32 && own64 >= 0 && own64 < cnt &&
                g64[own64] == (int64_t)q;
      for (int s = 0; ok && s < cnt; s++) ok = (size_t)g64[s] == g[s] && (size_t)g32[s] == g[s];
      wrong += !ok;
      bool same = offs[q + 1] - offs[q] == cnt;
      for (int s = 0; same && s < cnt; s++) same = (size_t)mem[offs[q] + s] == g[s];
      ggfx_diff += !same;
    }
    // the halo path of the N = 8 stage kernels
    size_t halo_wrong = 0;
    const bool hok = gf::box_halo8_ok(t);
    if (N == 8) {
      std::vector<double> gid(n);
      for (size_t q = 0; q < n; q++) gid[q] = (double)q;  // exact below 2^53
      std::vector<double> hl(gf::kHalo8);
      for (size_t e = 0; e < nel; e++) {
        int ec[3];
        gf::box_elem_coords(t, e, ec);
        const gf::BoxElemDeltas de = gf::box_elem_deltas(t, N, ec);
        const int64_t ebase = (int64_t)e * 512;
        for (double &h : hl) h = -1.0;
        // the element's own values as the kernels' 16-byte global-to-LDS
        // copy leaves them (the element's layout), then each lane's
        // partner values
        for (int q = 0; q < 512; q++) hl[q] = gid[ebase + q];
        for (int lane = 0; lane < 64; lane++) {
          gf::Halo8Rim<double> L;
          L.issue(gid.data(), de, ebase, lane);
          L.store(hl.data(), lane);
        }
        for (int lane = 0; lane < 64; lane++) {
          const int i = lane & 7, j = lane >> 3;
          const gf::Halo8Col hc(de, i, j);
          gf::static_for<0, 8>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const size_t q = (size_t)ebase + i + 8 * j + 64 * k;
            size_t g[8];
            const int cnt = gf::box_group(t, N, ec, q, i, j, k, g);
            int ad[8];
            const int hn = gf::halo8_order<k>(hc, de, ad);
            // the members in canonical order, absent partners at the zero
            // slot (which holds +0) in between
            int nm = 0;
            bool ok = hn == cnt && hl[gf::kHalo8Zero] == 0.0;
            for (int m = 0; ok && m < 8; m++) {
              if (ad[m] == gf::kHalo8Zero) continue;
              ok = nm < cnt && hl[ad[m]] == (double)g[nm];
              nm++;
            }
            ok = ok && nm == cnt;
            if (k != 0 && k != 7)
              for (int m = 4; ok && m < 8; m++) ok = ad[m] == gf::kHalo8Zero;
            halo_wrong += !ok;
          });
        }
      }
    }
    // the halo order must hold exactly where box_halo8_ok says it does (the
    // kernels use the halo only there)
    const bool halo_pass = N != 8 || (hok ? halo_wrong == 0 : halo_wrong > 0);
    const bool pass = wrong == 0 && halo_pass && (!c.expect_ggfx || ggfx_diff == 0);
    printf("p=%d ne=%d,%d,%d: n=%zu wrong=%zu ggfx_diff=%zu halo_ok=%d halo_wrong=%zu %s\n", c.p,
           c.ne[0], c.ne[1], c.ne[2], n, wrong, ggfx_diff, (int)hok, halo_wrong,
           pass ? "ok" : "FAIL");
    bad += !pass;
  }
  return bad ? 1 : 0;
}
