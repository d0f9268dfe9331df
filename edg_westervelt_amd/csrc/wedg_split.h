// Layout of the split tables of the two-lane stage kernels (wedg_stage2.h),
// shared by the host code that builds them.
#pragma once

#include <hip/hip_runtime.h>

namespace wedg {

constexpr int kPairWave = 32;   // elements per wave of the two-lane kernels

// LDS doubles per lane of a two-lane stage kernel: the lift vectors of the
// four sides (4 (2 RV + RU)), later the mass-matrix weights of the lane's rows.
__host__ __device__ constexpr int lds2_per_lane(int nu) { return nu >= 7 ? 48 : 36; }

// Offsets in the split table of one lane parity h (n_u = nu, n_v = nv in
// {nu, nu - 1}, nq = nu points, nh high-order points). A lane holds the rows
// a = 2k + h of a modal array, k < R = ceil(n / 2); a row a >= n (the parity-1
// lane's last row when n is odd) is a dummy whose entries are 0 in every
// table, so it stays 0. Point rows: p < nq / 2 (the lane's mirrored point),
// then the middle point (nq odd), whose entries of the other parity are 0.
struct SplitLayout {
    int RU, RV, NQS, NHH;
    int sPu, sSu, sPv, sSv, sPh;      // P_u, S_u, P_v, S_v (NQS x R), P_h (NHH x RV): [p * R + k]
    int sE1lu, sE1ru;                 // E1bl, E1br at the u rows (u traces; the b lifts)
    int sElv, sErv, sE1lv, sE1rv;     // Ebl, Ebr, E1bl, E1br at the v rows (v traces; the Rv lifts)
    int sQ, sD;                       // Q[2k'+h][2k+h] ([k' + RU k]); one-class u-solve factors ([k + RU jj])
    int n;
};

__host__ __device__ constexpr SplitLayout split_layout(int nu, int nv, int nh) {
    SplitLayout L{};
    L.RU = (nu + 1) / 2;
    L.RV = (nv + 1) / 2;
    L.NQS = (nu + 1) / 2;
    L.NHH = (nh + 1) / 2;
    const bool same = nv == nu;   // the v tables are the u tables
    int o = 0;
    L.sPu = o; o += L.NQS * L.RU;
    L.sSu = o; o += L.NQS * L.RU;
    if (same) { L.sPv = L.sPu; L.sSv = L.sSu; }
    else { L.sPv = o; o += L.NQS * L.RV; L.sSv = o; o += L.NQS * L.RV; }
    L.sPh = o; o += L.NHH * L.RV;
    L.sE1lu = o; o += L.RU;
    L.sE1ru = o; o += L.RU;
    L.sElv = o; o += L.RV;
    L.sErv = o; o += L.RV;
    if (same) { L.sE1lv = L.sE1lu; L.sE1rv = L.sE1ru; }
    else { L.sE1lv = o; o += L.RV; L.sE1rv = o; o += L.RV; }
    L.sQ = o; o += L.RU * L.RU;
    L.sD = o; o += L.RU * nu;
    L.n = o;
    return L;
}

}  // namespace wedg
