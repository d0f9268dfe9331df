// Layout of the split tables of the two-lane stage kernels (wedg_stage2.h),
// shared by the host code that builds them.
#pragma once

#include <hip/hip_runtime.h>

namespace wedg {

constexpr int kPairWave = 32;         // elements per wave of the two-lane kernels
constexpr int kLds2 = 36;              // LDS doubles per lane: lift vectors, then mass-matrix weights

// Offsets in the split table of one lane parity h (n_u = n_v = nu even, nh
// high-order points): values T[p][2k + h] of the 1D tables for the lane's
// parity, and the eigenvalue factors of a one-class spectral u-solve.
struct SplitLayout {
    int RU, NQH, NHH;
    int sP, sS, sPh, sEr, sE1r, sTEl, sTEr, sFE1l, sFE1r, sE1l, sQ, sD, n;
};

__host__ __device__ constexpr SplitLayout split_layout(int nu, int nh) {
    SplitLayout L{};
    L.RU = nu / 2;
    L.NQH = nu / 2;
    L.NHH = (nh + 1) / 2;
    int o = 0;
    L.sP = o; o += L.NQH * L.RU;     // P_u[p][2k+h], [p * RU + k]
    L.sS = o; o += L.NQH * L.RU;     // S_u[p][2k+h]
    L.sPh = o; o += L.NHH * L.RU;    // P_h[p][2k+h] (middle row: the odd columns are 0)
    L.sEr = o; o += L.RU;            // Ebr[2k+h]
    L.sE1r = o; o += L.RU;           // E1br[2k+h]
    L.sTEl = o; o += L.RU;           // (2/h) Ebl[2k+h]
    L.sTEr = o; o += L.RU;           // (2/h) Ebr
    L.sFE1l = o; o += L.RU;          // (4/h^2) E1bl
    L.sFE1r = o; o += L.RU;          // (4/h^2) E1br
    L.sE1l = o; o += L.RU;           // E1bl
    L.sQ = o; o += L.RU * L.RU;      // Q[2k'+h][2k+h], [k' + RU k]
    L.sD = o; o += L.RU * nu;        // one-class u-solve factors D[2k+h][jj], [k + RU jj]
    L.n = o;
    return L;
}

}  // namespace wedg
