// Device kernels of the EDG Westervelt right-hand side and RK4 stages (gfx950, FP64).
//
// One lane owns one element e = j + N k: it reads the element's modal
// coefficients straight from the state layout (N, N, n, n) — consecutive lanes
// read consecutive elements, so every load and store is coalesced — and keeps
// the element's small matrices in registers. The 1D Legendre tables live in a
// read-only buffer addressed only through compile-time offsets, so they are
// fetched with scalar loads and feed the FMAs as SGPR operands.
//
// Two kernels per right-hand side:
//   k_traces   the traces of every element on its four sides (reference
//              _traces!/_trace4!, src/rhs.jl:384-410, 573-580) into a trace
//              buffer, and with the global step-count policy the Chebyshev
//              intervals (src/rhs.jl:889-918) reduced to one count;
//   k_element  each element's own face slots from its traces and its
//              neighbours' (interior fluxes src/rhs.jl:620-700, boundary
//              fluxes src/rhs.jl:65-128, 738-765), the volume terms
//              (src/rhs.jl:317-333), the source (src/rhs.jl:340-350), the lifts
//              (src/rhs.jl:421-473), the u-equation (src/rhs.jl:361-382), the
//              Chebyshev mass solve (src/rhs.jl:889-939) and the RK4 stage
//              update (src/solver.jl:300-321) in its epilogue.
// Every face slot is computed by the element that owns it, so no slot has two
// writers and no floating-point atomics exist; results do not depend on the
// launch shape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wedg {

constexpr int kBlock = 256;      // lanes per block of the element kernels
constexpr int kMassBlock = 64;   // one wave per block for the mass solve (per-lane LDS)
constexpr int kMassLdsDoubles = 80;   // Chebyshev weights per lane kept in LDS: 4 waves fill 160 KB

template <int NU_, int NV_>
struct Dims {
    static constexpr int NU = NU_, NV = NV_, NQ = NU_;
    static constexpr int QU = NU_ - 1, QV = NV_ - 1;
    static constexpr int NH0 = (3 * QV + 1) / 2 + 1;                 // ceil(3 q_v / 2) + 1
    static constexpr int NH = NH0 > NU_ ? NH0 : NU_;                  // src/solver.jl:483
    static constexpr int NPK = (NU * NU * (NU * NU + 1)) / 2;
    // offsets in the table buffer
    static constexpr int oP = 0;                    // P_u  (NQ x NU), column-major
    static constexpr int oS = oP + NQ * NU;         // S_u  (NQ x NU)
    static constexpr int oW = oS + NQ * NU;         // w    (NQ)
    static constexpr int oEl = oW + NQ;             // Ebl  (NU)
    static constexpr int oEr = oEl + NU;            // Ebr
    static constexpr int oE1l = oEr + NU;           // E1bl
    static constexpr int oE1r = oE1l + NU;          // E1br
    static constexpr int oPh = oE1r + NU;           // P_h_v (NH x NV)
    static constexpr int oWh = oPh + NH * NV;       // W2d_h (NH x NH)
    static constexpr int nTab = oWh + NH * NH;
};

struct FaceArr {
    const double* p;   // [i][face] (face index fastest), or one value
    int uni;           // 1: the array is one constant
    int pad;
};

// Failure record of the device (reset by the host before each public call).
struct DevErr {
    int bkey;          // smallest _boundary_key (INT_MAX = none)
    int mfail;         // smallest failing element, 1-based (INT_MAX = none)
    int failed;        // serial of the stage that failed (0 = none)
    int nonfinite;     // first RK4 step with a non-finite entry (INT_MAX = none)
    int maxcheb;       // largest Chebyshev count taken
    int pad[3];
    unsigned long long cheb_wave_sum;   // sum over waves and stages of the count each wave ran
};

enum { F_HC2 = 0, F_HAF, F_HCF, F_HBF, F_VC2, F_VAF, F_VCF, F_VBF,
       F_HC2M, F_HAFM, F_VC2M, F_VAFM, F_HJC, F_VJC, F_COUNT };

struct KArgs {
    int N, E, stage, cheb_global;
    int serial, slot, one_sided, step;
    int kind[4];
    int nproj, pad0;
    double two_h, four_hh, theta, tau, lu, lv, omega, cheb_L;
    double ct, st, srct, cstage;
    double pc[5];                   // coefficients of the projections
    double gam[4], eta[4], sig[4], kb[4], kap[4];
    const double* tab;
    const int* cls_ca;              // class of (Cw, Aw, Ainv) per element, null = one class
    const double* Cw;               // per class (NQ*NQ), iy + NQ ix
    const double* Aw;
    const double* Ainv;             // per class (NU^2 x NU^2), symmetric, unpacked
    const int* cls_b;               // class of beta_hq per element (all 0 for one class)
    const double* beta;             // per class (NH*NH), p + NH q
    FaceArr f[F_COUNT];
    int fstride;                    // N (N + 1): stride of the point index of face arrays
    int pad1;
    const double* bcd[8];           // boundary data (NQ, N): sp, spn for bot, top, lft, rgt
    const double* proj[5];          // element-fastest [m][E]
    const double* yu; const double* yv;     // stage state Y_s
    const double* y0u; const double* y0v;   // base state y of the step
    double* accu; double* accv;             // RK4 accumulator
    double* outu; double* outv;             // Y_{s+1}, y (last stage) or (u_t, v_t)
    double* tr;                             // traces [arr][i][E]
    double* Rbuf;                           // right-hand side of the mass systems [m][E]
    double* Bbuf;                           // (c^2 grad phi_u, grad v_h) [m][E]
    DevErr* err;
    int* cnt;                               // Chebyshev counts, two slots
};

// value of a face array at flat index fi = i * (N (N + 1)) + face
__device__ __forceinline__ double fv(const FaceArr& a, long fi) { return a.p[a.uni ? 0 : fi]; }


// The pointer, opaque to the optimizer: loads through it are issued where they
// are used instead of being hoisted out of a loop and kept alive (the tables
// are small and read from the scalar cache at every use).
template <class T>
__device__ __forceinline__ const T* fresh(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(p));
#endif
    return p;
}

__device__ __forceinline__ bool skip_stage(const KArgs& a) {
    const int fl = a.err->failed;
    return fl != 0 && fl < a.serial;
}

// ---------------------------------------------------------------------------
// traces of one field on the four sides (src/rhs.jl:388-410): for side B
//   o[i] = sum_b P[i, b] sum_a El[a] X[a, b]
// and L/R the same on the transpose; sums in order, fused like muladd.
template <int NQ, int N, int oP, int oEl, int oEr>
__device__ __forceinline__ void trace4(const double* __restrict__ T, const double (&X)[N * N],
                                       double* __restrict__ tr, int arr0, int e, int E) {
    double oB[NQ], oT[NQ], oL[NQ], oR[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) { oB[i] = 0.0; oT[i] = 0.0; oL[i] = 0.0; oR[i] = 0.0; }
#pragma unroll
    for (int b = 0; b < N; ++b) {
        double sB = 0.0, sT = 0.0, sL = 0.0, sR = 0.0;
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const double x = X[a + N * b], xt = X[b + N * a];
            sB = fma(T[oEl + a], x, sB);
            sT = fma(T[oEr + a], x, sT);
            sL = fma(T[oEl + a], xt, sL);
            sR = fma(T[oEr + a], xt, sR);
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const double pp = T[oP + i + NQ * b];
            oB[i] = fma(pp, sB, oB[i]);
            oT[i] = fma(pp, sT, oT[i]);
            oL[i] = fma(pp, sL, oL[i]);
            oR[i] = fma(pp, sR, oR[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        tr[(long)((arr0 + 0) * NQ + i) * E + e] = oB[i];
        tr[(long)((arr0 + 1) * NQ + i) * E + e] = oT[i];
        tr[(long)((arr0 + 2) * NQ + i) * E + e] = oL[i];
        tr[(long)((arr0 + 3) * NQ + i) * E + e] = oR[i];
    }
}

// One row p of g = (2 beta) o (P_h V P_h^T) at the high-order points.
template <class D>
__device__ __forceinline__ void g_row(const double* __restrict__ T, const double (&V)[D::NV * D::NV],
                                      const double* __restrict__ beta, int p, double (&g)[D::NH]) {
    constexpr int NV = D::NV, NH = D::NH;
    double tp[NV];
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NV; ++a) s = fma(T[D::oPh + p + NH * a], V[a + NV * b], s);
        tp[b] = s;
    }
#pragma unroll
    for (int q = 0; q < NH; ++q) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NV; ++b) s = fma(tp[b], T[D::oPh + q + NH * b], s);
        g[q] = s * (2.0 * beta[p + NH * q]);
    }
}

// Interval data of the Chebyshev iteration from the extremes of g
// (src/rhs.jl:896-904): the a-priori step count, or 0 when 1 - 2 beta v_q <= 0
// somewhere (or a value is not a number).
__device__ __forceinline__ int cheb_interval(double gmin, double gmax, bool nan, double L, double& th,
                                             double& de, double& s1) {
    const double amin = 1.0 - gmax, amax = 1.0 - gmin;
    if (nan || !(amin > 0.0)) return 0;
    th = 0.5 * (amax + amin);
    de = fmax(0.5 * (amax - amin), 1e-15 * th);
    s1 = th / de;
    const int n = (int)ceil(L / acosh(s1));
    return n < 1 ? 1 : n;
}

// ---------------------------------------------------------------------------
// Traces of every element, and with the global policy the largest Chebyshev
// step count (src/rhs.jl:889-918) reduced over the grid.
template <int NU, int NV>
__global__ __launch_bounds__(kBlock) void k_traces(const KArgs a, const double* __restrict__ T0,
                                                   const double* __restrict__ beta0) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ, NH = D::NH;
    if (a.err->failed != 0) return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    const int E = a.E;
    int n_e = 1;
    if (e < E) {
        {
            double U[NU * NU];
#pragma unroll
            for (int m = 0; m < NU * NU; ++m) U[m] = a.yu[(long)m * E + e];
            trace4<NQ, NU, D::oP, D::oE1l, D::oE1r>(fresh(T0), U, a.tr, 0, e, E);
        }
        double V[NV * NV];
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) V[m] = a.yv[(long)m * E + e];
        trace4<NQ, NV, D::oP, D::oEl, D::oEr>(fresh(T0), V, a.tr, 4, e, E);
        trace4<NQ, NV, D::oP, D::oE1l, D::oE1r>(fresh(T0), V, a.tr, 8, e, E);
        if (a.cheb_global) {
            const double* beta = beta0 + (long)a.cls_b[e] * NH * NH;
            double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
            bool nan = false;
#pragma unroll 1
            for (int p = 0; p < NH; ++p) {
                double g[NH];
                g_row<D>(fresh(T0), V, beta, p, g);
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    gmin = fmin(gmin, g[q]);
                    gmax = fmax(gmax, g[q]);
                    nan = nan || (g[q] != g[q]);
                }
            }
            double th, de, s1;
            n_e = cheb_interval(gmin, gmax, nan, a.cheb_L, th, de, s1);
            if (n_e == 0) {
                atomicMin(&a.err->mfail, e + 1);
                n_e = 1;
            }
        }
    }
    if (a.cheb_global) {
        for (int off = 32; off > 0; off >>= 1) n_e = max(n_e, __shfl_xor(n_e, off));
        if ((threadIdx.x & 63) == 0) atomicMax(&a.cnt[a.slot], n_e);
    }
}

// ---------------------------------------------------------------------------
// Slot values (A, D, F) of one face point of an interior face, for the element
// on the "+" side (plus) or the "-" side (src/rhs.jl:620-700). vm, vp, ... are
// the traces of the "-" and "+" elements at the point; fi indexes the face
// arrays at the point.
__device__ __forceinline__ void interior_point(const KArgs& a, bool plus, bool hz, long fi, double vm, double vp,
                                               double dum, double dup, double dvm, double dvp, double& A,
                                               double& Dd, double& Ff) {
    const double th = a.theta, tau = a.tau, lu = a.lu, lv = a.lv, two_h = a.two_h;
    const FaceArr& fc2 = a.f[hz ? F_HC2 : F_VC2];
    const FaceArr& faf = a.f[hz ? F_HAF : F_VAF];
    if (a.one_sided) {
        const double c2m = fv(a.f[hz ? F_HC2M : F_VC2M], fi), c2p = fv(fc2, fi);
        const double afm = fv(a.f[hz ? F_HAFM : F_VAFM], fi), afp = fv(faf, fi);
        const double K = fv(a.f[hz ? F_HJC : F_VJC], fi);
        const double J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp) + K;
        const double G = two_h * ((1 - th) * (c2m * dum + afm * dvm) + th * (c2p * dup + afp * dvp));
        const double pen = (0.5 * (c2m + c2p)) * lu + (0.5 * (afm + afp)) * lv;
        const double s = vm - vp;
        if (plus) {
            const double dvsp = th * s - tau * J;
            A = pen * s - G - (1 - th) * K;
            Dd = -afp * dvsp;
            Ff = c2p * dvsp;
        } else {
            const double dvsm = (1 - th) * (vp - vm) - tau * J;
            A = G - pen * s - th * K;
            Dd = afm * dvsm;
            Ff = -c2m * dvsm;
        }
    } else {
        const double c2 = fv(fc2, fi), af = fv(faf, fi);
        const double jmp = vm - vp;
        const double gst = two_h * (c2 * (th * dup + (1 - th) * dum) + af * (th * dvp + (1 - th) * dvm));
        const double tj = tau * (c2 * (two_h * (dum - dup)) + af * (two_h * (dvm - dvp)));
        if (plus) {
            A = (c2 * lu + af * lv) * jmp - gst;
            Dd = -af * (th * jmp - tj);
            Ff = c2 * (th * jmp - tj);
        } else {
            A = gst - (c2 * lu + af * lv) * jmp;
            Dd = af * ((1 - th) * (vp - vm) - tj);
            Ff = -c2 * ((1 - th) * (vp - vm) - tj);
        }
    }
}

// Slot values of one point i of a physical boundary side s (src/rhs.jl:738-765);
// returns the point's failing _boundary_key (src/rhs.jl:732) or 0.
__device__ __forceinline__ int boundary_point(const KArgs& a, int s, long fi, int i, int m, int nq, double du,
                                              double vt, double dv, double& A, double& Dd, double& Ff) {
    const bool hz = s < 2;
    const double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    const double c2f = fv(a.f[hz ? F_HC2 : F_VC2], fi);
    const double af = fv(a.f[hz ? F_HAF : F_VAF], fi);
    const double cf = fv(a.f[hz ? F_HCF : F_VCF], fi);
    const double bf = fv(a.f[hz ? F_HBF : F_VBF], fi);
    const double two_h = a.two_h, om = a.omega, ct = a.ct, stt = a.st;
    const double gun = sgn * (two_h * du);
    const double gvn = sgn * (two_h * dv);
    const double spv = a.bcd[2 * s][i + nq * m], spnv = a.bcd[2 * s + 1][i + nq * m];
    const double gu_ex = sgn * (ct * spnv - spnv);
    const double gv_ex = sgn * (-om * stt * spnv);
    const double v_ex = -om * stt * spv;
    double F1, dvv;
    if (a.kind[s] == 3) {                                  // wall (src/rhs.jl:91-96)
        const double g2 = c2f * gu_ex + af * gv_ex;
        const double rho_w = c2f * gun + af * gvn - g2;
        F1 = g2;
        dvv = -(a.kap[s] / cf) * rho_w;
    } else {
        const double kb = a.kb[s];
        const int check = 1.0 - kb * bf * vt <= 0.0 ? 1 : (1.0 - kb * bf * v_ex <= 0.0 ? 2 : 0);
        if (check) {
            A = 0.0; Dd = 0.0; Ff = 0.0;
            return 2 * (s * nq * a.N + i + m * nq) + check;
        }
        const double g = a.gam[s], et = a.eta[s], sg = a.sig[s];   // src/rhs.jl:74-85
        const double rho_num = et * (c2f * gun + af * gvn) + g * cf * sqrt(1.0 - kb * bf * vt) * vt;
        const double rho_ex = et * (c2f * gu_ex + af * gv_ex) + g * cf * sqrt(1.0 - kb * bf * v_ex) * v_ex;
        const double rho_net = rho_num - rho_ex;
        const double sqf = sqrt(1.0 - kb * bf * vt);
        F1 = c2f * gun + af * gvn - (et + sg * g) * rho_net;
        dvv = -(g - sg * et) / (cf * sqf) * rho_net;
    }
    A = F1;
    Dd = sgn * (af * dvv);
    Ff = -sgn * (c2f * dvv);
    return 0;
}

// ---------------------------------------------------------------------------
// Volume terms (src/rhs.jl:317-333) scaled by -4/h^2 plus the source
// (src/rhs.jl:340-350) into R, and VB = (c^2 grad phi_u, grad v_h) into B.
// The point values are formed one point row p at a time and contracted back at
// once (rank-one updates), so only the modal arrays stay live.
template <class D, int n, bool UNIC>
__device__ __forceinline__ void vol_pass(const double* __restrict__ T0, const double* __restrict__ W0,
                                         const double* __restrict__ Wv, const double (&X)[n * n],
                                         int nout, double* __restrict__ out) {
    // out (nout x nout) += S^T (W o (S X P^T)) P + P^T (W o (P X S^T)) S over the points
    constexpr int NQ = D::NQ;
#pragma unroll 1
    for (int p = 0; p < NQ; ++p) {
        const double* __restrict__ T = fresh(T0);
        const double* __restrict__ W = UNIC ? fresh(W0) : Wv;
        double t1[n], t2[n];
#pragma unroll
        for (int b = 0; b < n; ++b) {
            double s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int aa = 0; aa < n; ++aa) {
                s1 = fma(T[D::oS + p + NQ * aa], X[aa + n * b], s1);
                s2 = fma(T[D::oP + p + NQ * aa], X[aa + n * b], s2);
            }
            t1[b] = s1; t2[b] = s2;
        }
        double z1[D::NU], z2[D::NU];
#pragma unroll
        for (int b = 0; b < D::NU; ++b) { z1[b] = 0.0; z2[b] = 0.0; }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double g1 = 0.0, g2 = 0.0;
#pragma unroll
            for (int b = 0; b < n; ++b) {
                g1 = fma(t1[b], T[D::oP + q + NQ * b], g1);
                g2 = fma(t2[b], T[D::oS + q + NQ * b], g2);
            }
            const double w = W[p + NQ * q];
            const double h1 = w * g1, h2 = w * g2;
#pragma unroll
            for (int b = 0; b < D::NU; ++b) {
                if (b < nout) {
                    z1[b] = fma(h1, T[D::oP + q + NQ * b], z1[b]);
                    z2[b] = fma(h2, T[D::oS + q + NQ * b], z2[b]);
                }
            }
        }
#pragma unroll
        for (int b = 0; b < D::NU; ++b)
#pragma unroll
            for (int aa = 0; aa < D::NU; ++aa)
                if (b < nout && aa < nout)
                    out[aa + nout * b] = fma(T[D::oS + p + NQ * aa], z1[b],
                                             fma(T[D::oP + p + NQ * aa], z2[b], out[aa + nout * b]));
    }
}

template <int NU, int NV, bool UNIC>
__global__ __launch_bounds__(kBlock) void k_volume(const KArgs a, const double* __restrict__ T0,
                                                   const double* __restrict__ Cw0, const double* __restrict__ Aw0,
                                                   double* __restrict__ R, double* __restrict__ B) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ;
    if (skip_stage(a)) return;
    const int E = a.E;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    const long cc = UNIC ? 0 : (long)a.cls_ca[e];
    const double* __restrict__ Cw = Cw0 + cc * NQ * NQ;
    const double* __restrict__ Aw = Aw0 + cc * NQ * NQ;
    double Rv[NV * NV];
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) Rv[m] = 0.0;
    {   // (c^2 grad phi_v, grad u_h)
        double U[NU * NU];
#pragma unroll
        for (int m = 0; m < NU * NU; ++m) U[m] = a.yu[(long)m * E + e];
        vol_pass<D, NU, UNIC>(T0, Cw0, Cw, U, NV, Rv);
    }
    double V[NV * NV];
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) V[m] = a.yv[(long)m * E + e];
    // (alpha grad phi_v, grad v_h)
    vol_pass<D, NV, UNIC>(T0, Aw0, Aw, V, NV, Rv);
    // scale and source, in the order of the reference's sum
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) {
        double r = (-a.four_hh) * Rv[m];
        for (int t = 0; t < a.nproj; ++t) r = r + a.pc[t] * a.proj[t][(long)m * E + e];
        R[(long)m * E + e] = r;
    }
    // (c^2 grad phi_u, grad v_h)
    double VB[NU * NU];
#pragma unroll
    for (int m = 0; m < NU * NU; ++m) VB[m] = 0.0;
    vol_pass<D, NV, UNIC>(T0, Cw0, Cw, V, NU, VB);
#pragma unroll
    for (int m = 0; m < NU * NU; ++m) B[(long)m * E + e] = VB[m];
}

// ---------------------------------------------------------------------------
// Face slots of the element's four sides from its own and its neighbours'
// traces, lifted into R and Fu (src/rhs.jl:421-473, 589-776); then the
// u-equation (src/rhs.jl:361-382) and the u part of the RK4 stage update
// (src/solver.jl:300-321).
template <class D>
struct SideGeom {
    int fidx, nb, m;
    bool bnd, hz, plus;
};

template <class D>
__device__ __forceinline__ SideGeom<D> side_geom(int s, int e, int j, int k, int N) {
    SideGeom<D> g;
    g.hz = s < 2;
    g.plus = (s == 0 || s == 2);
    if (s == 0)      { g.fidx = j + (N + 1) * k;     g.nb = e - 1; g.bnd = j == 0;     g.m = k; }
    else if (s == 1) { g.fidx = j + 1 + (N + 1) * k; g.nb = e + 1; g.bnd = j == N - 1; g.m = k; }
    else if (s == 2) { g.fidx = j + N * k;           g.nb = e - N; g.bnd = k == 0;     g.m = j; }
    else             { g.fidx = j + N * (k + 1);     g.nb = e + N; g.bnd = k == N - 1; g.m = j; }
    if (g.bnd) g.nb = e;    // not read on the boundary
    return g;
}

// Slot values (A, D, F) of point i of side s of element e and its failing
// boundary key (0 if none).
template <class D>
__device__ __forceinline__ int slot_point(const KArgs& a, const SideGeom<D>& g, int s, int i, int e, double& A,
                                          double& Dd, double& Ff) {
    constexpr int NQ = D::NQ;
    const int E = a.E;
    const double odu = a.tr[(long)((0 + s) * NQ + i) * E + e];
    const double ov = a.tr[(long)((4 + s) * NQ + i) * E + e];
    const double odv = a.tr[(long)((8 + s) * NQ + i) * E + e];
    const long fi = (long)i * a.fstride + g.fidx;
    if (g.bnd) return boundary_point(a, s, fi, i, g.m, NQ, odu, ov, odv, A, Dd, Ff);
    const int ns = s ^ 1;   // the neighbour's facing side: B <-> T, L <-> R
    const double ndu = a.tr[(long)((0 + ns) * NQ + i) * E + g.nb];
    const double nvv = a.tr[(long)((4 + ns) * NQ + i) * E + g.nb];
    const double ndv = a.tr[(long)((8 + ns) * NQ + i) * E + g.nb];
    if (g.plus) interior_point(a, true, g.hz, fi, nvv, ov, ndu, odu, ndv, odv, A, Dd, Ff);
    else interior_point(a, false, g.hz, fi, ov, nvv, odu, ndu, odv, ndv, A, Dd, Ff);
    return 0;
}

// Face slots of the element's four sides from its own and its neighbours'
// traces, lifted into R (src/rhs.jl:421-473, 589-776); then the same slots
// lifted into Fu, the u-equation (src/rhs.jl:361-382) and the u part of the
// RK4 stage update (src/solver.jl:300-321). The two lifts are separate passes
// over the sides so that R and Fu are never live together; the slot values,
// a few flops per point, are formed in both.
template <int NU, int NV, bool UNIC>
__global__ __launch_bounds__(kBlock) void k_faces(const KArgs a, const double* __restrict__ T0,
                                                  const double* __restrict__ Ai0, double* __restrict__ R,
                                                  const double* __restrict__ B) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ;
    if (skip_stage(a)) return;
    const int E = a.E, N = a.N;
    const int mfail_prev = a.cheb_global ? a.err->mfail : 0x7fffffff;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    const int j = e % N, k = e / N;
    int bkey = 0;
    {   // lifts of A and D into the v-equation (src/rhs.jl:449-463)
        double Rv[NV * NV];
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) Rv[m] = R[(long)m * E + e];
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const SideGeom<D> g = side_geom<D>(s, e, j, k, N);
            double YA[NV], YD[NV];
#pragma unroll
            for (int b = 0; b < NV; ++b) { YA[b] = 0.0; YD[b] = 0.0; }
#pragma unroll 1
            for (int i = 0; i < NQ; ++i) {
                const double* __restrict__ T = fresh(T0);
                double A, Dd, Ff;
                const int k2 = slot_point<D>(a, g, s, i, e, A, Dd, Ff);
                if (k2) bkey = bkey == 0 ? k2 : min(bkey, k2);
                // Y_Z[b] += P[i, b] (w[i] Z[i]), summed in point order (src/rhs.jl:434-441)
                const double wA = T[D::oW + i] * A, wD = T[D::oW + i] * Dd;
#pragma unroll
                for (int b = 0; b < NV; ++b) {
                    YA[b] = fma(T[D::oP + i + NQ * b], wA, YA[b]);
                    YD[b] = fma(T[D::oP + i + NQ * b], wD, YD[b]);
                }
            }
            const double* __restrict__ T = fresh(T0);
            const int oE = g.plus ? D::oEl : D::oEr;
            const int oE1 = g.plus ? D::oE1l : D::oE1r;
#pragma unroll
            for (int b = 0; b < NV; ++b)
#pragma unroll
                for (int aa = 0; aa < NV; ++aa) {
                    double r = Rv[aa + NV * b];
                    if (g.hz) {
                        r = r + (a.two_h * T[oE + aa]) * YA[b];
                        r = r + (a.four_hh * T[oE1 + aa]) * YD[b];
                    } else {
                        r = r + (a.two_h * YA[aa]) * T[oE + b];
                        r = r + (a.four_hh * YD[aa]) * T[oE1 + b];
                    }
                    Rv[aa + NV * b] = r;
                }
        }
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) R[(long)m * E + e] = Rv[m];
    }
    if (bkey) {
        atomicMin(&a.err->bkey, bkey);
        a.err->failed = a.serial;
    }
    if (mfail_prev != 0x7fffffff) a.err->failed = a.serial;

    constexpr int n = NU * NU;
    double b[n];   // b = VB - Fu, then b[1] += V[1] (src/rhs.jl:366-370)
    {
        double Fu[n];
#pragma unroll
        for (int m = 0; m < n; ++m) Fu[m] = 0.0;
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            const SideGeom<D> g = side_geom<D>(s, e, j, k, N);
            double YF[NU];
#pragma unroll
            for (int bb = 0; bb < NU; ++bb) YF[bb] = 0.0;
#pragma unroll 1
            for (int i = 0; i < NQ; ++i) {
                const double* __restrict__ T = fresh(T0);
                double A, Dd, Ff;
                slot_point<D>(a, g, s, i, e, A, Dd, Ff);
                const double wF = T[D::oW + i] * Ff;
#pragma unroll
                for (int bb = 0; bb < NU; ++bb) YF[bb] = fma(T[D::oP + i + NQ * bb], wF, YF[bb]);
            }
            const double* __restrict__ T = fresh(T0);
            const int oE1 = g.plus ? D::oE1l : D::oE1r;
#pragma unroll
            for (int bb = 0; bb < NU; ++bb)
#pragma unroll
                for (int aa = 0; aa < NU; ++aa) {
                    if (g.hz) Fu[aa + NU * bb] = Fu[aa + NU * bb] + (1.0 * T[oE1 + aa]) * YF[bb];
                    else Fu[aa + NU * bb] = Fu[aa + NU * bb] + (1.0 * YF[aa]) * T[oE1 + bb];
                }
        }
#pragma unroll
        for (int i = 0; i < n; ++i) b[i] = B[(long)i * E + e] - Fu[i];
    }
    b[0] += a.yv[e];

    // u-equation u_t = A_e^-1 b (src/rhs.jl:361-382). The inverse of the class is
    // stored unpacked; it is symmetric, so row i is the contiguous column i and
    // each output is finished (and written) before the next one starts.
    const long cc = UNIC ? 0 : (long)a.cls_ca[e];
    const double c = a.cstage;
    bool fin = true;
#pragma unroll 1
    for (int mm = 0; mm < n; ++mm) {
        const double* __restrict__ Ar = (UNIC ? fresh(Ai0) : Ai0 + cc * (n * n)) + (long)n * mm;
        double kv = 0.0;
#pragma unroll
        for (int jj = 0; jj < n; ++jj) kv = fma(Ar[jj], b[jj], kv);
        const long idx = (long)mm * E + e;
        if (a.stage == 0) {
            a.outu[idx] = kv;
        } else if (a.stage == 1) {
            a.outu[idx] = a.y0u[idx] + c * kv;
            a.accu[idx] = kv;
        } else if (a.stage < 4) {
            a.outu[idx] = a.y0u[idx] + c * kv;
            a.accu[idx] = a.accu[idx] + 2.0 * kv;
        } else {
            const double y = a.y0u[idx] + c * (a.accu[idx] + kv);
            a.outu[idx] = y;
            fin = fin && isfinite(y);
        }
    }
    if (!fin) atomicMin(&a.err->nonfinite, a.step);
}

// ---------------------------------------------------------------------------
// One point row p of K d = P_h^T (G o (P_h d P_h^T)) P_h accumulated into r.
template <class D>
__device__ __forceinline__ void kd_row(const double* __restrict__ T, int p, const double (&d)[D::NV * D::NV],
                                       const double (&Gp)[D::NH], double (&r)[D::NV * D::NV]) {
    constexpr int NV = D::NV, NH = D::NH;
    double tp[NV], zp[NV];
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        double s = 0.0;
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) s = fma(T[D::oPh + p + NH * aa], d[aa + NV * b], s);
        tp[b] = s;
        zp[b] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < NH; ++q) {
        double y = 0.0;
#pragma unroll
        for (int b = 0; b < NV; ++b) y = fma(tp[b], T[D::oPh + q + NH * b], y);
        y *= Gp[q];
#pragma unroll
        for (int b = 0; b < NV; ++b) zp[b] = fma(y, T[D::oPh + q + NH * b], zp[b]);
    }
#pragma unroll
    for (int b = 0; b < NV; ++b)
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) r[aa + NV * b] = fma(T[D::oPh + p + NH * aa], zp[b], r[aa + NV * b]);
}

// The mass system M(v_h) X = R by Chebyshev iteration (src/rhs.jl:886-939)
// and the v part of the RK4 stage update. One wave per block: the first rows
// of the weights G = 2 beta v_q W of every lane live in LDS ([entry][lane],
// conflict-free), the remaining rows in registers.
template <int NU, int NV>
__global__ __launch_bounds__(kMassBlock) void k_mass(const KArgs a, const double* __restrict__ T0,
                                                     const double* __restrict__ beta0,
                                                     const double* __restrict__ R) {
    using D = Dims<NU, NV>;
    constexpr int NH = D::NH;
    constexpr int RL0 = kMassLdsDoubles / NH;
    constexpr int RL = RL0 < NH ? RL0 : NH;          // rows of G in LDS
    constexpr int RT = NH - RL;                      // rows of G in registers
    __shared__ double Gs[RL * NH * kMassBlock];
    if (skip_stage(a)) return;
    const int E = a.E;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.cnt[a.slot ^ 1] = 1;   // the next stage's count
    const int lane = threadIdx.x;
    const int e = blockIdx.x * kMassBlock + lane;
    int ncheb = 0;
    if (e < E) {
        double Gt[RT > 0 ? RT * NH : 1];
        double th = 1.0, de = 1.0, s1 = 1.0;
        int n_e;
        {
            double V[NV * NV];
#pragma unroll
            for (int m = 0; m < NV * NV; ++m) V[m] = a.yv[(long)m * E + e];
            const double* beta = beta0 + (long)a.cls_b[e] * NH * NH;
            double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
            bool nan = false;
#pragma unroll 1
            for (int p = 0; p < RL; ++p) {
                const double* __restrict__ T = fresh(T0);
                double g[NH];
                g_row<D>(T, V, beta, p, g);
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    gmin = fmin(gmin, g[q]);
                    gmax = fmax(gmax, g[q]);
                    nan = nan || (g[q] != g[q]);
                    Gs[(p * NH + q) * kMassBlock + lane] = g[q] * T[D::oWh + p + NH * q];
                }
            }
#pragma unroll
            for (int pt = 0; pt < RT; ++pt) {
                const double* __restrict__ T = fresh(T0);
                const int p = RL + pt;
                double g[NH];
                g_row<D>(T, V, beta, p, g);
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    gmin = fmin(gmin, g[q]);
                    gmax = fmax(gmax, g[q]);
                    nan = nan || (g[q] != g[q]);
                    Gt[pt * NH + q] = g[q] * T[D::oWh + p + NH * q];
                }
            }
            n_e = cheb_interval(gmin, gmax, nan, a.cheb_L, th, de, s1);
        }
        double X[NV * NV];
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) X[m] = 0.0;
        if (n_e == 0) {
            if (!a.cheb_global) {
                atomicMin(&a.err->mfail, e + 1);
                a.err->failed = a.serial;
            }
        } else {
            const int n = a.cheb_global ? a.cnt[a.slot] : n_e;
            ncheb = n;
            double r[NV * NV], d[NV * NV];
#pragma unroll
            for (int m = 0; m < NV * NV; ++m) { r[m] = R[(long)m * E + e]; d[m] = r[m] / th; }
            double rho = 1.0 / s1;
            for (int step = 1; step <= n; ++step) {
#pragma unroll
                for (int m = 0; m < NV * NV; ++m) X[m] += d[m];
                if (step == n) break;
                // r <- (r - d) + K d
#pragma unroll
                for (int m = 0; m < NV * NV; ++m) r[m] = r[m] - d[m];
#pragma unroll 1
                for (int p = 0; p < RL; ++p) {
                    double Gp[NH];
#pragma unroll
                    for (int q = 0; q < NH; ++q) Gp[q] = Gs[(p * NH + q) * kMassBlock + lane];
                    kd_row<D>(fresh(T0), p, d, Gp, r);
                }
#pragma unroll
                for (int pt = 0; pt < RT; ++pt) {
                    double Gp[NH];
#pragma unroll
                    for (int q = 0; q < NH; ++q) Gp[q] = Gt[pt * NH + q];
                    kd_row<D>(fresh(T0), RL + pt, d, Gp, r);
                }
                const double rhon = 1.0 / (2.0 * s1 - rho);
                const double c1 = rhon * rho, c2 = 2.0 * rhon / de;
#pragma unroll
                for (int m = 0; m < NV * NV; ++m) d[m] = c1 * d[m] + c2 * r[m];
                rho = rhon;
            }
        }
        // RK4 stage update, v part
        const double c = a.cstage;
        bool fin = true;
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) {
            const long idx = (long)m * E + e;
            const double kv = X[m];
            if (a.stage == 0) {
                a.outv[idx] = kv;
            } else if (a.stage == 1) {
                a.outv[idx] = a.y0v[idx] + c * kv;
                a.accv[idx] = kv;
            } else if (a.stage < 4) {
                a.outv[idx] = a.y0v[idx] + c * kv;
                a.accv[idx] = a.accv[idx] + 2.0 * kv;
            } else {
                const double y = a.y0v[idx] + c * (a.accv[idx] + kv);
                a.outv[idx] = y;
                fin = fin && isfinite(y);
            }
        }
        if (!fin) atomicMin(&a.err->nonfinite, a.step);
    }
    for (int off = 32; off > 0; off >>= 1) ncheb = max(ncheb, __shfl_xor(ncheb, off));
    if (lane == 0 && ncheb > 0) {
        atomicMax(&a.err->maxcheb, ncheb);
        atomicAdd(&a.err->cheb_wave_sum, (unsigned long long)ncheb);
    }
}

}  // namespace wedg
