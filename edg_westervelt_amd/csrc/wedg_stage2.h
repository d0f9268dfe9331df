// The RK stage with two lanes per element (gfx950, FP64): k_prep2 / k_stage2.
//
// The one-lane kernels (wedg_kernels.h) give each element one lane and keep
// all of its small matrices in that lane's registers: at q = 5 that is the
// whole register file, so a SIMD runs one wave, and with 64 elements per wave
// a mesh of up to 256 x 256 elements has at most one wave per SIMD — every
// memory phase of the stage stalls the SIMD, and a mesh that fills only part
// of the chip takes a whole wave's latency per stage. Here the lanes i and
// i + 32 of a wave share element i: a wave holds 32 elements, the grid has
// twice the waves and each lane about half the registers, so a stage's
// latency is about half a one-lane wave's work, and at q <= 5 two waves share
// a SIMD. Each half-wave reads 32 consecutive elements of one mode (256
// contiguous bytes), so the element-fastest state is read in whole cache
// lines as by the one-lane kernel.
//
// How an element is shared (n_q = n_u; n_v = n_u or n_u - 1; h = lane / 32,
// the lane's parity):
//  * a modal array X[a + n b] (a the first index) is split by the parity of
//    a: lane h holds the rows a = 2k + h, k < R = ceil(n/2), as X[k + R b];
//    with n odd the parity-1 lane's last row is a dummy: it is read as 0
//    (its loads and stores fall outside the buffer's range and the hardware
//    returns 0 and drops the store), and every table that adds into a row
//    has 0 there, so it stays 0;
//  * point values along a mirrored point index p (face points, rows of
//    volume and mass-matrix points) are split by halves: lane 0 holds
//    p < NR/2, lane 1 the mirror point NR - 1 - p in the same slot; with NR
//    odd both hold the middle point;
//  * a contraction over the split first index: each lane forms the partial
//    sum of its parity with its half of the 1D table (the split table, kept in
//    the wave's LDS), the pair exchange their partial sums (v_permlane32_swap:
//    afterwards every lane holds both the parity-0 and the parity-1 value) and
//    each keeps its own mirrored point: the 1D tables are values
//    or derivatives of Legendre polynomials at symmetric points,
//    T[NR-1-p][a] = (-1)^a T[p][a] (values), -(-1)^a T[p][a] (derivatives), so
//    the even and odd partial sums give both mirrored points, and only one
//    parity reaches a middle point; the way back (points to modes) exchanges
//    the point values and each lane accumulates the modes of its parity;
//  * contractions over the second index run on each lane's own rows with the
//    uniform tables (SGPR operands), as in the one-lane kernel;
//  * point data of a mirrored row (quadrature weights, and the coefficient
//    records of a one-class problem) are read from row p by both lanes: the
//    host admits a problem only when those records are mirror-symmetric to
//    rounding (1e-13 relative); records of per-element classes are read from
//    each lane's own row.
// The results differ from the one-lane kernel's by rounding only (sums are
// grouped differently). Everything else — the failure records, the count
// slots of the global Chebyshev count, the RK4 registers — is as in k_stage,
// with the count slots per wave of 32 elements.
#pragma once

#include "wedg_kernels.h"
#include "wedg_split.h"

namespace wedg {

static_assert(kPairWave * 2 == kWave, "a lane pair per element");

// The per-lane sizes of one (n_u, n_v) pair.
template <class D>
struct Dims2 {
    static constexpr int NU = D::NU, NV = D::NV, NQ = D::NQ, NH = D::NH;
    static constexpr int RU = (NU + 1) / 2, RV = (NV + 1) / 2;   // rows per lane
    static constexpr int NQS = (NQ + 1) / 2;                     // point slots per lane (pairs, middle)
    static constexpr int nU = RU * NU, nV = RV * NV;             // a lane's share of a modal array
    static constexpr SplitLayout SL = split_layout(NU, NV, NH);
    static constexpr int NLDS = lds2_per_lane(NU);
};

// Both values of a lane pair: every lane gets the value of the pair's parity-0
// lane (a) and of its parity-1 lane (b) (lanes i and i + 32; v_permlane32_swap
// exchanges the upper half of its first operand with the lower half of its
// second).
struct Pair {
    double a, b;
};

__device__ __forceinline__ Pair xboth(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    Pair r;
    r.a = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
    r.b = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
    return r;
#else
    return Pair{x, x};
#endif
}

__device__ __forceinline__ int xany(int x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const auto r = __builtin_amdgcn_permlane32_swap((unsigned)x, (unsigned)x, false, false);
    return (int)(r[0] | r[1]);
#else
    return x;
#endif
}

// A lane's view of its element: parity, the sign of the mirrored half, the
// byte offset of its rows in element-fastest arrays, its split table in LDS.
struct Lane2 {
    int h;                // lane parity (lane / 32)
    double sg;            // 1 (parity 0), -1 (parity 1)
    int voffE;            // (h E + e) * 8: rows 2k + h of element e in an [m][E] array
    const double* t2;     // the lane's split table (LDS)
};

// A byte offset past the buffer range of the element-fastest arrays (their
// records cover 0x7fffffff bytes): a load there returns 0, a store is dropped.
constexpr unsigned kDummyOffset = 0x80000000u;

#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ double ld_rs(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void st_rs(__amdgpu_buffer_rsrc_t r, int voff, int soff, double x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(wedg_u32x2, x), r, voff, soff, 0);
}
#endif

// The lane's rows 2k + h of an element's n x n modal array (element-fastest
// [m][E]; voff = (h E + e) * 8); the dummy row (n odd, parity 1) reads 0.
template <int n>
__device__ __forceinline__ void eload2(const double* p, int E, int voff, int h, double (&x)[((n + 1) / 2) * n]) {
    constexpr int R = (n + 1) / 2;
    const int vlast = ((n & 1) && h) ? (int)kDummyOffset : voff;
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t r = ersrc(p);
#pragma unroll
    for (int b = 0; b < n; ++b)
#pragma unroll
        for (int k = 0; k < R; ++k) x[k + R * b] = ld_rs(r, k == R - 1 ? vlast : voff, (2 * k + n * b) * E * 8);
#else
    for (int b = 0; b < n; ++b)
        for (int k = 0; k < R; ++k)
            x[k + R * b] = (k == R - 1 && vlast != voff) ? 0.0 : p[(long)(2 * k + n * b) * E + voff / 8];
#endif
}

template <int n>
__device__ __forceinline__ void estore2(double* p, int E, int voff, int h, const double (&x)[((n + 1) / 2) * n]) {
    constexpr int R = (n + 1) / 2;
    const int vlast = ((n & 1) && h) ? (int)kDummyOffset : voff;
#if defined(__HIP_DEVICE_COMPILE__)
    const __amdgpu_buffer_rsrc_t r = ersrc(p);
#pragma unroll
    for (int b = 0; b < n; ++b)
#pragma unroll
        for (int k = 0; k < R; ++k) st_rs(r, k == R - 1 ? vlast : voff, (2 * k + n * b) * E * 8, x[k + R * b]);
#else
    for (int b = 0; b < n; ++b)
        for (int k = 0; k < R; ++k)
            if (!(k == R - 1 && vlast != voff)) p[(long)(2 * k + n * b) * E + voff / 8] = x[k + R * b];
#endif
}

// First element of the block (one wave of 32 elements), blocks dealt to the
// XCDs in contiguous bands as in block_first_element; in a batched launch, the
// block's place in its instance's range.
template <bool BT>
__device__ __forceinline__ int block_first_element2() {
    if constexpr (BT) {
        return ((int)blockIdx.x - ((cint*)bargs()->bfirst)[block_instance()]) * kPairWave;
    } else {
        const int b = blockIdx.x, nb = gridDim.x;
        const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
        return lb * kPairWave;
    }
}

// ---------------------------------------------------------------------------
// Pair combinations. After a contraction over the split index the pair's
// partial sums are ev (parity 0, x.a) and od (parity 1, x.b):
//   lane 0 (point p):        ev + od
//   lane 1 (point NR-1-p):   ev - od (values), od - ev (derivatives)
__device__ __forceinline__ double comb_fwd(Pair x, double sg, bool der) {
    return der ? fma(sg, x.a, x.b) : fma(sg, x.b, x.a);
}
// The way back from the point values y[p] (lane 0, x.a) and y[NR-1-p] (lane 1,
// x.b): lane 0 accumulates the even modes, lane 1 the odd ones:
//   values:      even modes y[p] + y[NR-1-p], odd modes y[p] - y[NR-1-p]
//   derivatives: even modes y[p] - y[NR-1-p], odd modes y[p] + y[NR-1-p]
__device__ __forceinline__ double comb_back(Pair x, double sg, bool der) {
    return der ? fma(-sg, x.b, x.a) : fma(sg, x.b, x.a);
}

// Point values (own mirrored points p < NR/2) of the vector x over the second
// (unsplit) index with the uniform table T (NR x n, column-major): lane 0 the
// point p, lane 1 the point NR-1-p.
template <int NR, int n, class X>
__device__ __forceinline__ double pt_unsplit(cdbl* T, int p, X x, double sg) {
    double ev = 0.0, od = 0.0;
#pragma unroll
    for (int a = 0; a < n; a += 2) ev = fma(T[p + NR * a], x(a), ev);
#pragma unroll
    for (int a = 1; a < n; a += 2) od = fma(T[p + NR * a], x(a), od);
    return fma(sg, od, ev);
}

// The point index of the lane's point slot p.
template <int NQ>
__device__ __forceinline__ int point_of(int p, int h) {
    return p < NQ / 2 ? (h ? NQ - 1 - p : p) : NQ / 2;
}

// ---------------------------------------------------------------------------
// Traces at the lane's point slots of one side (0 B, 1 T: the first index at
// its ends; 2 L, 3 R: the second index), o[p] for p < NQS (src/rhs.jl:384-410).
// VF: the field is v (n_v modes), else u; DER: the derivative row E1.
template <class D, bool VF, bool DER, int sd>
__device__ __forceinline__ void side_trace2(const KArgs& A, const Lane2& L,
                                            const double (&X)[(VF ? Dims2<D>::nV : Dims2<D>::nU)],
                                            double (&o)[Dims2<D>::NQS]) {
    using Q2 = Dims2<D>;
    constexpr int NQ = D::NQ, n = VF ? D::NV : D::NU, R = (n + 1) / 2;
    constexpr SplitLayout SL = Q2::SL;
    static_assert(VF || DER, "the u traces are derivatives");
    cdbl* T = ctab(A.tab);
    if (sd < 2) {
        // sv[b] = sum_a E[a] X[a, b]: the lane's parity, then the pair's sum
        constexpr int oE = DER ? (VF ? SL.sE1rv : SL.sE1ru) : SL.sErv;
        double sv[n];
#pragma unroll
        for (int b = 0; b < n; ++b) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) s = fma(L.t2[oE + k], X[k + R * b], s);
            const Pair x = xboth(s);
            // + end: ev + od; - end: ev - od (values), od - ev (derivatives)
            sv[b] = sd == 1 ? x.a + x.b : (DER ? x.b - x.a : x.a - x.b);
        }
#pragma unroll
        for (int p = 0; p < NQ / 2; ++p) o[p] = pt_unsplit<NQ, n>(T + D::oP, p, [&](int b) { return sv[b]; }, L.sg);
        if (NQ & 1) o[NQ / 2] = row_mid<NQ, n, false>(T + D::oP, [&](int b) { return sv[b]; });
    } else {
        // sv[a] = sum_b E[b] X[a, b] on the lane's rows, then the points over a
        constexpr int oEr = DER ? D::oE1r : D::oEr;
        constexpr int oP = VF ? SL.sPv : SL.sPu;
        double sv[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double ev = 0.0, od = 0.0;
#pragma unroll
            for (int b = 0; b < n; b += 2) ev = fma(T[oEr + b], X[k + R * b], ev);
#pragma unroll
            for (int b = 1; b < n; b += 2) od = fma(T[oEr + b], X[k + R * b], od);
            sv[k] = sd == 3 ? ev + od : (DER ? od - ev : ev - od);
        }
#pragma unroll
        for (int p = 0; p < Q2::NQS; ++p) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < R; ++k) s = fma(L.t2[oP + p * R + k], sv[k], s);
            const Pair x = xboth(s);
            o[p] = p < NQ / 2 ? comb_fwd(x, L.sg, false) : x.a + x.b;   // the middle point: even modes only
        }
    }
}

// Modes 2k + h of sum_i P[i, m] y[i] (m < n of the field, VF: v) from the
// lane's point slots y (the lift vectors, src/rhs.jl:434-441).
template <class D, bool VF>
__device__ __forceinline__ void lift_vec2(const Lane2& L, const double (&y)[Dims2<D>::NQS],
                                          double (&z)[VF ? Dims2<D>::RV : Dims2<D>::RU]) {
    using Q2 = Dims2<D>;
    constexpr int NQ = D::NQ, R = VF ? Q2::RV : Q2::RU;
    constexpr int oP = VF ? Q2::SL.sPv : Q2::SL.sPu;
    double c[Q2::NQS];
#pragma unroll
    for (int p = 0; p < NQ / 2; ++p) c[p] = comb_back(xboth(y[p]), L.sg, false);
    if (NQ & 1) c[NQ / 2] = y[NQ / 2];   // the middle point: the table is 0 at the odd modes
#pragma unroll
    for (int k = 0; k < R; ++k) {
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < Q2::NQS; ++p) s = fma(L.t2[oP + p * R + k], c[p], s);
        z[k] = s;
    }
}

// The slots of the element's physical-boundary sides at the lane's point
// slots (src/rhs.jl:738-765), with the weight checks (kind 0, the stage's
// serial). The middle point (n_q odd) is formed by both lanes alike.
template <class D, bool UNI>
__device__ __forceinline__ void stage_boundary2(const Lane2& L, const double (&U)[Dims2<D>::nU],
                                                const double (&V)[Dims2<D>::nV], int e, int serial, double ct,
                                                double st) {
    using Q2 = Dims2<D>;
    constexpr int NQ = D::NQ, NQS = Q2::NQS;
    CKArgs& B = *kargs<D::BT>();
    const int N = B.N, j = e % N, k = e / N;
    int bkey = 0;
    auto side = [&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const bool bnd = s == 0 ? j == 0 : s == 1 ? j == N - 1 : s == 2 ? k == 0 : k == N - 1;
        if (!bnd) return;
        CKArgs& A = *kargs<D::BT>();
        double du[NQS], vt[NQS], dv[NQS];
        side_trace2<D, false, true, s>(A, L, U, du);
        side_trace2<D, true, false, s>(A, L, V, vt);
        side_trace2<D, true, true, s>(A, L, V, dv);
        const int m = s < 2 ? k : j;
        const int fidx = s == 0 ? (N + 1) * k : s == 1 ? N + (N + 1) * k : s == 2 ? j : j + N * N;
        double* bs = A.bsl + (long)s * NQ * 3 * N + m;
        BPoint pd[NQS];
#pragma unroll
        for (int p = 0; p < NQS; ++p) {
            const int i = point_of<NQ>(p, L.h);
            pd[p] = boundary_data<UNI, (s < 2)>(A, s, (long)i * A.fstride + fidx, i, m, NQ);
        }
#pragma unroll
        for (int p = 0; p < NQS; ++p) {
            const int i = point_of<NQ>(p, L.h);
            double Av, Dd, Ff;
            const int k2 = boundary_point<UNI, (s < 2)>(A, s, pd[p], i, m, NQ, du[p], vt[p], dv[p], ct, st, Av, Dd, Ff);
            if (k2) bkey = bkey == 0 ? k2 : min(bkey, k2);
            bs[(long)(i * 3 + 0) * N] = Av;
            bs[(long)(i * 3 + 1) * N] = Dd;
            bs[(long)(i * 3 + 2) * N] = Ff;
        }
    };
    side(std::integral_constant<int, 0>());
    side(std::integral_constant<int, 1>());
    side(std::integral_constant<int, 2>());
    side(std::integral_constant<int, 3>());
    if (bkey) atomicMin(&B.err->fkey, fail_key(serial, 0, bkey));
}

// Face slots of side s at the lane's point slots (interior fluxes
// src/rhs.jl:620-700 with the neighbour's facing traces recomputed from its
// state; on a physical boundary the slots stage_boundary2 formed), reduced to
// the lane's share of the side's lift vectors Y_A, Y_D (v modes), Y_F (u
// modes) of its parity, into its LDS slots.
template <class D, bool UNI, int s>
__device__ __forceinline__ void face_side2(const Lane2& L, int e, int j, int k, int N, int E,
                                           const double (&U)[Dims2<D>::nU], const double (&V)[Dims2<D>::nV],
                                           double* Ys, int lane) {
    using Q2 = Dims2<D>;
    constexpr int NQ = D::NQ, NV = D::NV, NU = D::NU, NQS = Q2::NQS, RU = Q2::RU, RV = Q2::RV;
    constexpr bool hz = s < 2, plus = (s == 0 || s == 2);
    constexpr int ns = s ^ 1;
    CKArgs& A = *kargs<D::BT>();
    int fidx, nb, m;
    bool bnd;
    if (s == 0)      { fidx = j + (N + 1) * k;     nb = e - 1; bnd = j == 0;     m = k; }
    else if (s == 1) { fidx = j + 1 + (N + 1) * k; nb = e + 1; bnd = j == N - 1; m = k; }
    else if (s == 2) { fidx = j + N * k;           nb = e - N; bnd = k == 0;     m = j; }
    else             { fidx = j + N * (k + 1);     nb = e + N; bnd = k == N - 1; m = j; }
    if (bnd) nb = e;
    double odu[NQS], ov[NQS], odv[NQS], ndu[NQS], nvv[NQS], ndv[NQS];
    side_trace2<D, false, true, s>(A, L, U, odu);
    side_trace2<D, true, false, s>(A, L, V, ov);
    side_trace2<D, true, true, s>(A, L, V, odv);
    const int voffN = (L.h * E + nb) * 8;
    {
        double X[Q2::nU];
        eload2<NU>(A.yu, E, voffN, L.h, X);
        side_trace2<D, false, true, ns>(A, L, X, ndu);
    }
    {
        double X[Q2::nV];
        eload2<NV>(A.yv, E, voffN, L.h, X);
        side_trace2<D, true, false, ns>(A, L, X, nvv);
        side_trace2<D, true, true, ns>(A, L, X, ndv);
    }
    double wA[NQS], wD[NQS], wF[NQS];
    {
        CKArgs& I = *kargs<D::BT>();
        cdbl* Tw = ctab(I.tab) + D::oW;   // w[NQ-1-p] = w[p] (admitted to 1e-13)
#pragma unroll
        for (int p = 0; p < NQS; ++p) {
            const int i = point_of<NQ>(p, L.h);
            const long fi = (long)i * I.fstride + fidx;
            double Av, Dd, Ff;
            if (plus) interior_point<UNI>(I, true, hz, fi, nvv[p], ov[p], ndu[p], odu[p], ndv[p], odv[p], Av, Dd, Ff);
            else interior_point<UNI>(I, false, hz, fi, ov[p], nvv[p], odu[p], ndu[p], odv[p], ndv[p], Av, Dd, Ff);
            wA[p] = Tw[p] * Av;
            wD[p] = Tw[p] * Dd;
            wF[p] = Tw[p] * Ff;
        }
    }
    if (bnd) {
        CKArgs& B = *kargs<D::BT>();
        cdbl* Tw = ctab(B.tab) + D::oW;
        const double* bs = B.bsl + (long)s * NQ * 3 * N + m;
#pragma unroll
        for (int p = 0; p < NQS; ++p) {
            const int i = point_of<NQ>(p, L.h);
            wA[p] = Tw[p] * bs[(long)(i * 3 + 0) * N];
            wD[p] = Tw[p] * bs[(long)(i * 3 + 1) * N];
            wF[p] = Tw[p] * bs[(long)(i * 3 + 2) * N];
        }
    }
    double YA[RV], YD[RV], YF[RU];
    lift_vec2<D, true>(L, wA, YA);
    lift_vec2<D, true>(L, wD, YD);
    lift_vec2<D, false>(L, wF, YF);
    constexpr int y0 = s * (2 * RV + RU);
#pragma unroll
    for (int b = 0; b < RV; ++b) {
        Ys[(y0 + b) * kWave + lane] = YA[b];
        Ys[(y0 + RV + b) * kWave + lane] = YD[b];
    }
#pragma unroll
    for (int b = 0; b < RU; ++b) Ys[(y0 + 2 * RV + b) * kWave + lane] = YF[b];
}

// The lifts of side s (src/rhs.jl:449-473) on the lane's rows: Rv += (2/h) E (x) Y_A
// + (4/h^2) E1 (x) Y_D, b -= E1 (x) Y_F on a horizontal side (the lift vectors'
// modes of both parities from the pair's LDS slots), the factors transposed on
// a vertical side (the lane's own modes).
template <class D, int s>
__device__ __forceinline__ void apply_lifts2(const Lane2& L, const double* Ys, int lane, double (&Rv)[Dims2<D>::nV],
                                             double (&bv)[Dims2<D>::nU]) {
    using Q2 = Dims2<D>;
    constexpr int NV = D::NV, NU = D::NU, RV = Q2::RV, RU = Q2::RU;
    constexpr SplitLayout SL = Q2::SL;
    constexpr bool hz = s < 2, plus = (s == 0 || s == 2);
    constexpr int y0 = s * (2 * RV + RU);
    CKArgs& A = *kargs<D::BT>();
    if (hz) {
        const int l0 = lane & 31, l1 = l0 + 32;
        const double two_h = A.two_h, four_hh = A.four_hh;
        constexpr int oE = plus ? SL.sElv : SL.sErv, oE1 = plus ? SL.sE1lv : SL.sE1rv;
        constexpr int oF1 = plus ? SL.sE1lu : SL.sE1ru;
        double te[RV], te1[RV];
#pragma unroll
        for (int k = 0; k < RV; ++k) {
            te[k] = two_h * L.t2[oE + k];       // (2/h) E (src/rhs.jl:449-463)
            te1[k] = four_hh * L.t2[oE1 + k];   // (4/h^2) E1
        }
#pragma unroll
        for (int b = 0; b < NV; ++b) {
            const int src = (b & 1) ? l1 : l0;
            const double ya = Ys[(y0 + b / 2) * kWave + src], yd = Ys[(y0 + RV + b / 2) * kWave + src];
#pragma unroll
            for (int k = 0; k < RV; ++k) {
                double r = Rv[k + RV * b];
                r = fma(te[k], ya, r);
                r = fma(te1[k], yd, r);
                Rv[k + RV * b] = r;
            }
        }
#pragma unroll
        for (int b = 0; b < NU; ++b) {
            const int src = (b & 1) ? l1 : l0;
            const double yf = Ys[(y0 + 2 * RV + b / 2) * kWave + src];
#pragma unroll
            for (int k = 0; k < RU; ++k) bv[k + RU * b] = fma(-L.t2[oF1 + k], yf, bv[k + RU * b]);
        }
    } else {
        cdbl* T = ctab(A.tab);
        constexpr int oE = plus ? D::oTEl : D::oTEr, oE1 = plus ? D::oFE1l : D::oFE1r, oF1 = plus ? D::oE1l : D::oE1r;
        double YA[RV], YD[RV], YF[RU];
#pragma unroll
        for (int k = 0; k < RV; ++k) {
            YA[k] = Ys[(y0 + k) * kWave + lane];
            YD[k] = Ys[(y0 + RV + k) * kWave + lane];
        }
#pragma unroll
        for (int k = 0; k < RU; ++k) YF[k] = Ys[(y0 + 2 * RV + k) * kWave + lane];
#pragma unroll
        for (int b = 0; b < NV; ++b)
#pragma unroll
            for (int k = 0; k < RV; ++k) {
                double r = Rv[k + RV * b];
                r = fma(YA[k], T[oE + b], r);
                r = fma(YD[k], T[oE1 + b], r);
                Rv[k + RV * b] = r;
            }
#pragma unroll
        for (int b = 0; b < NU; ++b)
#pragma unroll
            for (int k = 0; k < RU; ++k) bv[k + RU * b] = fma(-YF[k], T[oF1 + b], bv[k + RU * b]);
    }
}

// ---------------------------------------------------------------------------
// Volume terms (src/rhs.jl:317-333) on the lane's share:
//   out += S^T (W o (S X P^T)) P + P^T (W o (P X S^T)) S
// for X of field VI (n modes) and out of field VO (NO modes), one mirrored
// pair of point rows (p, NQ-1-p) at a time, lane 0 row p and lane 1 row
// NQ-1-p; the middle row (p = NQ/2, NQ odd) in both lanes. DX: the x-term (S on
// the first index).
template <class D, bool VI, bool VO, bool DX>
__device__ __forceinline__ void vol_term2(const KArgs& A, const Lane2& L, int p,
                                          const double (&X)[(VI ? Dims2<D>::nV : Dims2<D>::nU)], const double (&w)[D::NQ],
                                          double (&out)[(VO ? Dims2<D>::nV : Dims2<D>::nU)]) {
    using Q2 = Dims2<D>;
    constexpr int NQ = D::NQ, n = VI ? D::NV : D::NU, R = (n + 1) / 2;
    constexpr int NO = VO ? D::NV : D::NU, RO = (NO + 1) / 2;
    constexpr SplitLayout SL = Q2::SL;
    const bool mid = (NQ & 1) && p == NQ / 2;
    const int oI = (DX ? (VI ? SL.sSv : SL.sSu) : (VI ? SL.sPv : SL.sPu)) + p * R;    // row p of A, input rows
    const int oO = (DX ? (VO ? SL.sSv : SL.sSu) : (VO ? SL.sPv : SL.sPu)) + p * RO;   // output rows
    cdbl* B = ctab(A.tab) + (DX ? D::oP : D::oS);   // the second index, uniform
    double t[n];
#pragma unroll
    for (int b = 0; b < n; ++b) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < R; ++k) s = fma(L.t2[oI + k], X[k + R * b], s);
        const Pair x = xboth(s);
        t[b] = mid ? x.a + x.b : comb_fwd(x, L.sg, DX);   // the middle row: one parity is 0
    }
    double z[NO];
    {
        double g[NQ];
        rows_sym<NQ, n, !DX>(B, [&](int b) { return t[b]; }, g);
#pragma unroll
        for (int q = 0; q < NQ; ++q) g[q] = w[q] * g[q];
        cols_sym<NQ, NO, !DX>(B, g, z);
    }
#pragma unroll
    for (int b = 0; b < NO; ++b) {
        const double c = mid ? z[b] : comb_back(xboth(z[b]), L.sg, DX);
#pragma unroll
        for (int k = 0; k < RO; ++k) out[k + RO * b] = fma(L.t2[oO + k], c, out[k + RO * b]);
    }
}

// The lane's weight row W[prow + NQ q]: for one class (UNI) row p in both
// lanes (uniform, the record is mirror-symmetric), otherwise the lane's own
// mirrored row prow of its element's class.
template <class D, bool UNI, bool VI, bool VO>
__device__ __forceinline__ void vol_pass2(const Lane2& L, const double (&X)[(VI ? Dims2<D>::nV : Dims2<D>::nU)],
                                          const double* Wsrc, double (&out)[(VO ? Dims2<D>::nV : Dims2<D>::nU)]) {
    constexpr int NQ = D::NQ;
#pragma unroll 1
    for (int p = 0; p < (NQ + 1) / 2; ++p) {
        double w[NQ];
        if (UNI) {
            cdbl* Wr = ctab(Wsrc) + p;
#pragma unroll
            for (int q = 0; q < NQ; ++q) w[q] = Wr[NQ * q];
        } else {
            const int prow = point_of<NQ>(p, L.h);
            const double* Wr = vfresh(Wsrc + prow);
#pragma unroll
            for (int q = 0; q < NQ; ++q) w[q] = Wr[NQ * q];
        }
        vol_term2<D, VI, VO, true>(*kargs<D::BT>(), L, p, X, w, out);
        __builtin_amdgcn_sched_barrier(0);
        vol_term2<D, VI, VO, false>(*kargs<D::BT>(), L, p, X, w, out);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// ---------------------------------------------------------------------------
// The values of 2 beta v_h at the lane's rows of high-order points and the
// mass-matrix weights G = 2 beta v_q W_h (src/rhs.jl:889-904): for each own row
// p < (NH+1)/2 (the middle one, NH odd, in both lanes) the callback
// f(p, q, g, Gv) with g = 2 beta v_q and Gv = g W_h. W_h and a one-class beta
// record are read from row p by both lanes (mirror-symmetric).
template <class D, bool UNI, class F>
__device__ __forceinline__ void mass_rows2(const KArgs& A, const Lane2& L, const double (&V)[Dims2<D>::nV],
                                           int cls_b, F f) {
    using Q2 = Dims2<D>;
    constexpr int NV = D::NV, NH = D::NH, RV = Q2::RV;
    constexpr SplitLayout SL = Q2::SL;
    const double* bq = A.beta + (UNI ? 0 : (long)cls_b * NH * NH);
#pragma unroll
    for (int p = 0; p < SL.NHH; ++p) {
        const bool mid = (NH & 1) && p == NH / 2;
        const int prow = (mid || !L.h) ? p : NH - 1 - p;
        double t[NV];
#pragma unroll
        for (int b = 0; b < NV; ++b) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < RV; ++k) s = fma(L.t2[SL.sPh + p * RV + k], V[k + RV * b], s);
            t[b] = comb_fwd(xboth(s), L.sg, false);
        }
        cdbl* T = ctab(A.tab);
        double y[NH];
        rows_sym<NH, NV, false>(T + D::oPh, [&](int b) { return t[b]; }, y);
        const double* br = vfresh(bq + prow);
        cdbl* bu = ctab(A.beta) + p;
#pragma unroll
        for (int q = 0; q < NH; ++q) {
            const double g = y[q] * (2.0 * (UNI ? bu[NH * q] : br[NH * q]));
            f(p, q, g, g * T[D::oWh + p + NH * q]);
        }
    }
}

// The element's a-priori Chebyshev count of the state V (0: 1 - 2 beta v_h <= 0
// somewhere, or not a number), the pair's extremes combined.
template <class D, bool UNI>
__device__ __forceinline__ int cheb_count2(const KArgs& A, const Lane2& L, const double (&V)[Dims2<D>::nV],
                                           int cls_b, double& th, double& de, double& s1) {
    double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
    bool nan = false;
    mass_rows2<D, UNI>(A, L, V, cls_b, [&](int, int, double g, double) {
        gmin = fmin(gmin, g);
        gmax = fmax(gmax, g);
        nan = nan || (g != g);
    });
    const Pair mn = xboth(gmin), mx = xboth(gmax);
    return cheb_interval(fmin(mn.a, mn.b), fmax(mx.a, mx.b), xany(nan ? 1 : 0) != 0, A.cheb_L, th, de, s1);
}

// Writes the wave's largest count into the per-wave slot (waves of 32 elements).
template <class D>
__device__ __forceinline__ void wave_count2(int n_e, int slot, int E) {
    for (int off = 32; off > 0; off >>= 1) n_e = max(n_e, __shfl_xor(n_e, off));
    if ((threadIdx.x & 63) == 0) {
        const int nw = (E + kPairWave - 1) / kPairWave;
        kargs<D::BT>()->cnt[slot * nw + block_first_element2<D::BT>() / kPairWave] = n_e;
    }
}

// The lane's parity and element, and the wave's copy of the split table in
// LDS (both parities, n doubles each; one wave per block, so the copy is
// visible to the wave after its own stores).
template <class D>
__device__ __forceinline__ Lane2 lane2_setup(int ec, int E, double* T2s) {
    constexpr SplitLayout SL = Dims2<D>::SL;
    const int lane = threadIdx.x;
    const double* t2 = kargs<D::BT>()->tab2;
    for (int i = lane; i < 2 * SL.n; i += kWave) T2s[i] = t2[i];
    Lane2 L;
    L.h = lane >> 5;
    L.sg = L.h ? -1.0 : 1.0;
    L.voffE = (L.h * E + ec) * 8;
    L.t2 = T2s + L.h * SL.n;
    return L;
}

// The first stage of a call: the Chebyshev counts of its input state.
template <int NU, int NV, int MODE, bool BT = false>
__global__ __launch_bounds__(kWave) void k_prep2(const std::conditional_t<BT, BatchArgs, KArgs> args) {
    (void)args;
    using D = Dims<NU, NV, BT>;
    using Q2 = Dims2<D>;
    constexpr bool UNI = MODE == MODE_UNIFORM;
    __shared__ double T2s[2 * Q2::SL.n];
    const int E = kargs<D::BT>()->E;
    const int e = block_first_element2<D::BT>() + (threadIdx.x & 31);
    const int ec = e < E ? e : E - 1;
    const bool global = kargs<D::BT>()->cheb_global != 0;
    const Lane2 L = lane2_setup<D>(ec, E, T2s);
    double V[Q2::nV];
    eload2<NV>(kargs<D::BT>()->yv, E, L.voffE, L.h, V);
    int n_e = 1;
    if (global) {
        double th, de, s1;
        const int cb = UNI ? 0 : kargs<D::BT>()->cls_b[ec];
        n_e = cheb_count2<D, UNI>(*kargs<D::BT>(), L, V, cb, th, de, s1);
        if (n_e == 0) {
            if (e < E && L.h == 0) atomicMin(&kargs<D::BT>()->err->fkey, fail_key(kargs<D::BT>()->serial, 1, e + 1));
            n_e = 1;
        }
        if (e >= E) n_e = 1;
        wave_count2<D>(n_e, kargs<D::BT>()->slot, E);
    }
}

// RK4 stage update of the lane's rows of an nn x nn modal array
// (src/solver.jl:300-321), as rk_store.
template <int nn>
__device__ __forceinline__ bool rk_store2(int stage, double c, const double (&kv)[((nn + 1) / 2) * nn],
                                          const double (&y)[((nn + 1) / 2) * nn],
                                          const double (&ac)[((nn + 1) / 2) * nn], double* acc, double* out, int E,
                                          int voffE, int h) {
    constexpr int n = ((nn + 1) / 2) * nn;
    bool fin = true;
    double o[n];
    auto st = [&](double* p, const double (&x)[n]) { estore2<nn>(p, E, voffE, h, x); };
    if (stage == 0) {
        st(out, kv);
    } else if (stage == 1) {
#pragma unroll
        for (int m = 0; m < n; ++m) o[m] = y[m] + c * kv[m];
        st(out, o);
        st(acc, kv);
    } else if (stage < 4) {
#pragma unroll
        for (int m = 0; m < n; ++m) o[m] = y[m] + c * kv[m];
        st(out, o);
#pragma unroll
        for (int m = 0; m < n; ++m) o[m] = ac[m] + 2.0 * kv[m];
        st(acc, o);
    } else {
#pragma unroll
        for (int m = 0; m < n; ++m) {
            o[m] = y[m] + c * (ac[m] + kv[m]);
            fin = fin && isfinite(o[m]);
        }
        st(out, o);
    }
    return fin;
}

// One RK stage, two lanes per element (k_stage with the element's work shared
// by the pair; MODE_UNIFORM or MODE_SEPARABLE: the spectral u-solve). Two
// waves per SIMD up to n_u = 6, one at n_u = 7.
template <int NU, int NV, int MODE, bool BT = false>
__global__ __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(NU >= 7 ? 1 : 2))) void k_stage2(
    const std::conditional_t<BT, BatchArgs, KArgs> args) {
    (void)args;
    using D = Dims<NU, NV, BT>;
    using Q2 = Dims2<D>;
    static_assert(MODE == MODE_UNIFORM || MODE == MODE_SEPARABLE, "spectral u-solve modes only");
    constexpr int NQ = D::NQ, NH = D::NH, RU = Q2::RU, RV = Q2::RV;
    constexpr bool UNI = MODE == MODE_UNIFORM;
    constexpr int nU = Q2::nU, nV = Q2::nV;
    constexpr SplitLayout SL = Q2::SL;
    constexpr int NLDS = Q2::NLDS;
    constexpr int NP = NH / 2;                            // mirrored pairs of high-order rows
    constexpr int PL = NLDS / NH < NP ? NLDS / NH : NP;   // pair rows of G in LDS
    constexpr int PR = NP - PL;
    static_assert(4 * (2 * RV + RU) <= NLDS, "lift vectors fit the LDS slots");
    __shared__ double Gs[NLDS * kWave];
    __shared__ double T2s[2 * SL.n];
    if (skip_stage(*kargs<D::BT>())) return;
    const int E = kargs<D::BT>()->E, N = kargs<D::BT>()->N;
    const bool global = kargs<D::BT>()->cheb_global != 0;
    const int lane = threadIdx.x;
    const int e = block_first_element2<D::BT>() + (lane & 31);
    const bool live = e < E;
    const int ec = live ? e : E - 1;
    const int nw = (E + kPairWave - 1) / kPairWave;
    const Lane2 L = lane2_setup<D>(ec, E, T2s);
    int nglob = 1;
    double U[nU], V[nV];
    {
        int m = 1;
        if (global) {
            const int* wc = kargs<D::BT>()->cnt + kargs<D::BT>()->slot * nw;
            for (int i0 = lane; i0 < nw; i0 += 8 * kWave) {
                int t[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) t[u] = i0 + u * kWave < nw ? wc[i0 + u * kWave] : 1;
#pragma unroll
                for (int u = 0; u < 8; ++u) m = max(m, t[u]);
            }
        }
        eload2<NU>(kargs<D::BT>()->yu, E, L.voffE, L.h, U);
        eload2<NV>(kargs<D::BT>()->yv, E, L.voffE, L.h, V);
        if (global) {
            for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
            nglob = m;
        }
    }
    const int j = ec % N, k = ec / N;
    if (live && (j == 0 || j == N - 1 || k == 0 || k == N - 1)) {
        CKArgs& A = *kargs<D::BT>();
        stage_boundary2<D, UNI>(L, U, V, ec, A.serial, A.ct, A.st);
    }
    const long cc = UNI ? 0 : (long)kargs<D::BT>()->cls_ca[ec];
    const double* Wc = kargs<D::BT>()->Cw + cc * NQ * NQ;
    const double* Wa = kargs<D::BT>()->Aw + cc * NQ * NQ;

    // ---- face slots and the lane's share of the lift vectors, into LDS
    face_side2<D, UNI, 0>(L, ec, j, k, N, E, U, V, Gs, lane);
    phase_barrier();
    face_side2<D, UNI, 1>(L, ec, j, k, N, E, U, V, Gs, lane);
    phase_barrier();
    face_side2<D, UNI, 2>(L, ec, j, k, N, E, U, V, Gs, lane);
    phase_barrier();
    face_side2<D, UNI, 3>(L, ec, j, k, N, E, U, V, Gs, lane);
    phase_barrier();

    // ---- volume terms
    double Rv[nV], bv[nU];
#pragma unroll
    for (int m = 0; m < nV; ++m) Rv[m] = 0.0;
#pragma unroll
    for (int m = 0; m < nU; ++m) bv[m] = 0.0;
    const double v0 = L.h == 0 ? V[0] : 0.0;   // the (0, 0) mode lives in lane 0
    vol_pass2<D, UNI, true, false>(L, V, Wc, bv);
    vol_pass2<D, UNI, true, true>(L, V, Wa, Rv);
    phase_barrier();
    vol_pass2<D, UNI, false, true>(L, U, Wc, Rv);
    phase_barrier();
    {
        const double s = -kargs<D::BT>()->four_hh;
#pragma unroll
        for (int m = 0; m < nV; ++m) Rv[m] = s * Rv[m];
        const int np = kargs<D::BT>()->nproj;
        for (int t = 0; t < np; ++t) {
            const double pc = kargs<D::BT>()->pc[t];
            double pv[nV];
            eload2<NV>(kargs<D::BT>()->proj[t], E, L.voffE, L.h, pv);
#pragma unroll
            for (int m = 0; m < nV; ++m) Rv[m] = Rv[m] + pc * pv[m];
        }
    }
    apply_lifts2<D, 0>(L, Gs, lane, Rv, bv);
    apply_lifts2<D, 1>(L, Gs, lane, Rv, bv);
    apply_lifts2<D, 2>(L, Gs, lane, Rv, bv);
    apply_lifts2<D, 3>(L, Gs, lane, Rv, bv);
    phase_barrier();
    bv[0] += v0;

    // ---- u-equation by the spectral solve (parity blocks: no exchange)
    double ku[nU];
    {
        cdbl* Q = ctab(kargs<D::BT>()->tab) + D::oQ;
        const double* qs = L.t2 + SL.sQ;
        double Y1[nU];
        // Y1[i][b] = sum_{aa of i's parity} Q[aa][i] b[aa][b]
#pragma unroll
        for (int b = 0; b < NU; ++b)
#pragma unroll
            for (int kk = 0; kk < RU; ++kk) {
                double s = 0.0;
#pragma unroll
                for (int k2 = 0; k2 < RU; ++k2) s = fma(qs[k2 + RU * kk], bv[k2 + RU * b], s);
                Y1[kk + RU * b] = s;
            }
        // the eigenvalue factors D[i][jj] of the lane's rows i = 2kk + h: the
        // split table's copy for one class, the class record otherwise (the
        // dummy row reads row n_u - 1 of the record and multiplies 0)
        const double* Dv = vfresh(kargs<D::BT>()->Ainv + cc * NU * NU);
#pragma unroll
        for (int kk = 0; kk < RU; ++kk)
#pragma unroll
            for (int jj = 0; jj < NU; ++jj) {
                double s = 0.0;
#pragma unroll
                for (int b = (jj & 1); b < NU; b += 2) s = fma(Y1[kk + RU * b], Q[b + NU * jj], s);
                const int di = min(2 * kk + L.h, NU - 1);
                bv[kk + RU * jj] = s * (UNI ? L.t2[SL.sD + kk + RU * jj] : Dv[di + NU * jj]);
            }
#pragma unroll
        for (int jj = 0; jj < NU; ++jj)
#pragma unroll
            for (int kk = 0; kk < RU; ++kk) {
                double s = 0.0;
#pragma unroll
                for (int k2 = 0; k2 < RU; ++k2) s = fma(qs[kk + RU * k2], bv[k2 + RU * jj], s);
                Y1[kk + RU * jj] = s;
            }
#pragma unroll
        for (int b = 0; b < NU; ++b)
#pragma unroll
            for (int kk = 0; kk < RU; ++kk) {
                double s = 0.0;
#pragma unroll
                for (int jj = (b & 1); jj < NU; jj += 2) s = fma(Y1[kk + RU * jj], Q[b + NU * jj], s);
                ku[kk + RU * b] = s;
            }
    }
    {
        CKArgs& A = *kargs<D::BT>();
        if (live && A.recu) estore2<NU>(A.recu, E, L.voffE, L.h, ku);
        double yu0[nU], acu[nU];
        if (A.stage >= 1) eload2<NU>(A.y0u, E, L.voffE, L.h, yu0);
        if (A.stage >= 2) eload2<NU>(A.accu, E, L.voffE, L.h, acu);
        if (live && !rk_store2<NU>(A.stage, A.cstage, ku, yu0, acu, A.accu, A.outu, E, L.voffE, L.h)) {
            atomicMin(&kargs<D::BT>()->err->nonfinite, kargs<D::BT>()->step);
            atomicMin(&kargs<D::BT>()->err->nf_serial, kargs<D::BT>()->serial);
        }
    }
    phase_barrier();

    // ---- the mass system M(v_h) X = Rv by Chebyshev iteration (src/rhs.jl:886-939)
    double Gt[PR > 0 ? PR * NH : 1];
    double Gm[NH];
    double th = 1.0, de = 1.0, s1 = 1.0;
    int n_e;
    const int cb = UNI ? 0 : kargs<D::BT>()->cls_b[ec];
    {
        double Vr[nV];
        eload2<NV>(kargs<D::BT>()->yv, E, L.voffE, L.h, Vr);
        double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
        bool nan = false;
        mass_rows2<D, UNI>(*kargs<D::BT>(), L, Vr, cb, [&](int p, int q, double g, double Gv) {
            gmin = fmin(gmin, g);
            gmax = fmax(gmax, g);
            nan = nan || (g != g);
            if ((NH & 1) && p == NH / 2) Gm[q] = Gv;
            else if (p < PL) Gs[(p * NH + q) * kWave + lane] = Gv;
            else Gt[(p - PL) * NH + q] = Gv;
        });
        const Pair mn = xboth(gmin), mx = xboth(gmax);
        n_e = cheb_interval(fmin(mn.a, mn.b), fmax(mx.a, mx.b), xany(nan ? 1 : 0) != 0, kargs<D::BT>()->cheb_L, th,
                            de, s1);
    }
    phase_barrier();
    double X[nV];
#pragma unroll
    for (int m = 0; m < nV; ++m) X[m] = 0.0;
    int ncheb = 0;
    if (n_e == 0) {
        if (!global && live && L.h == 0)
            atomicMin(&kargs<D::BT>()->err->fkey, fail_key(kargs<D::BT>()->serial, 1, e + 1));
    } else {
        const int n = global ? nglob : n_e;
        ncheb = n;
        double r[nV], d[nV];
#pragma unroll
        for (int m = 0; m < nV; ++m) { r[m] = Rv[m]; d[m] = r[m] / th; }
        double rho = 1.0 / s1;
        cdbl* Ph = ctab(kargs<D::BT>()->tab) + D::oPh;
        for (int step = 1; step <= n; ++step) {
#pragma unroll
            for (int m = 0; m < nV; ++m) X[m] += d[m];
            if (step == n) break;
#pragma unroll
            for (int m = 0; m < nV; ++m) r[m] = r[m] - d[m];
            // r += K d, one own point row p (lane 0: p, lane 1: NH-1-p) at a time
#pragma unroll
            for (int p = 0; p < SL.NHH; ++p) {
                const bool mid = (NH & 1) && p == NH / 2;
                const double* tph = L.t2 + SL.sPh + p * RV;
                double t[D::NV];
#pragma unroll
                for (int b = 0; b < D::NV; ++b) {
                    double s = 0.0;
#pragma unroll
                    for (int kk = 0; kk < RV; ++kk) s = fma(tph[kk], d[kk + RV * b], s);
                    t[b] = comb_fwd(xboth(s), L.sg, false);
                }
                double y[NH];
                rows_sym<NH, D::NV, false>(Ph, [&](int b) { return t[b]; }, y);
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    const double G = mid ? Gm[q] : (p < PL ? Gs[(p * NH + q) * kWave + lane] : Gt[(p - PL) * NH + q]);
                    y[q] *= G;
                }
                double z[D::NV];
                cols_sym<NH, D::NV, false>(Ph, y, z);
#pragma unroll
                for (int b = 0; b < D::NV; ++b) {
                    const double c = mid ? z[b] : comb_back(xboth(z[b]), L.sg, false);
#pragma unroll
                    for (int kk = 0; kk < RV; ++kk) r[kk + RV * b] = fma(tph[kk], c, r[kk + RV * b]);
                }
            }
            const double rhon = 1.0 / (2.0 * s1 - rho);
            const double c1 = rhon * rho, c2 = 2.0 * rhon / de;
#pragma unroll
            for (int m = 0; m < nV; ++m) d[m] = c1 * d[m] + c2 * r[m];
            rho = rhon;
        }
    }
    int n_next = 1;
    {
        CKArgs& A = *kargs<D::BT>();
        double yv0[nV], acv[nV];
        const int stage = A.stage;
        const double c = A.cstage;
        if (live && A.recv) estore2<NV>(A.recv, E, L.voffE, L.h, X);
        if (stage >= 1) eload2<NV>(A.y0v, E, L.voffE, L.h, yv0);
        if (stage >= 2) eload2<NV>(A.accv, E, L.voffE, L.h, acv);
        if (live && !rk_store2<NV>(stage, c, X, yv0, acv, A.accv, A.outv, E, L.voffE, L.h)) {
            atomicMin(&kargs<D::BT>()->err->nonfinite, kargs<D::BT>()->step);
            atomicMin(&kargs<D::BT>()->err->nf_serial, kargs<D::BT>()->serial);
        }
        if (A.has_next) {
            double Vn[nV];
#pragma unroll
            for (int m = 0; m < nV; ++m) Vn[m] = stage < 4 ? yv0[m] + c * X[m] : yv0[m] + c * (acv[m] + X[m]);
            if (global) {
                double t1, t2, t3;
                n_next = cheb_count2<D, UNI>(*kargs<D::BT>(), L, Vn, cb, t1, t2, t3);
                if (n_next == 0) {
                    if (live && L.h == 0)
                        atomicMin(&kargs<D::BT>()->err->fkey, fail_key(kargs<D::BT>()->serial + 1, 1, e + 1));
                    n_next = 1;
                }
                if (!live) n_next = 1;
            }
        }
    }
    if (global && kargs<D::BT>()->has_next) wave_count2<D>(n_next, kargs<D::BT>()->slot ^ 1, E);
    for (int off = 32; off > 0; off >>= 1) ncheb = max(ncheb, __shfl_xor(ncheb, off));
    if (lane == 0 && ncheb > 0) {
        unsigned* ws = kargs<D::BT>()->wstat;
        const int w = block_first_element2<D::BT>() / kPairWave;
        ws[w] += (unsigned)ncheb;
        ws[nw + w] = max(ws[nw + w], (unsigned)ncheb);
    }
}

}  // namespace wedg
