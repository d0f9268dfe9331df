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

constexpr int kBlock = 256;

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

__device__ __forceinline__ double fv(const FaceArr& a, int i, int idx, int stride) {
    return a.uni ? a.p[0] : a.p[(long)i * stride + idx];
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

// G = (2 beta) o (P_h V P_h^T) and the interval data of the Chebyshev
// iteration (src/rhs.jl:889-907). Returns the a-priori step count, or 0 when
// 1 - 2 beta v_q <= 0 somewhere. G is left without the weights W.
template <class D>
__device__ __forceinline__ int cheb_setup(const double* __restrict__ T, const double (&V)[D::NV * D::NV],
                                          const double* __restrict__ beta, double L,
                                          double (&G)[D::NH * D::NH], double& th, double& de, double& s1) {
    constexpr int NV = D::NV, NH = D::NH;
    double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
#pragma unroll
    for (int p = 0; p < NH; ++p) {
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
            const double g = s * (2.0 * beta[p + NH * q]);
            G[p + NH * q] = g;
            gmin = fmin(gmin, g);
            gmax = fmax(gmax, g);
        }
    }
    const double amin = 1.0 - gmax, amax = 1.0 - gmin;
    if (!(amin > 0.0)) return 0;
    th = 0.5 * (amax + amin);
    de = fmax(0.5 * (amax - amin), 1e-15 * th);
    s1 = th / de;
    int n = (int)ceil(L / acosh(s1));
    return n < 1 ? 1 : n;
}

// ---------------------------------------------------------------------------
template <int NU, int NV>
__global__ __launch_bounds__(kBlock) void k_traces(const KArgs a) {
    using D = Dims<NU, NV>;
    if (a.err->failed != 0) return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    const int E = a.E;
    int n_e = 1;
    if (e < E) {
        const double* __restrict__ T = a.tab;
        double U[NU * NU], V[NV * NV];
#pragma unroll
        for (int m = 0; m < NU * NU; ++m) U[m] = a.yu[(long)m * E + e];
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) V[m] = a.yv[(long)m * E + e];
        trace4<D::NQ, NU, D::oP, D::oE1l, D::oE1r>(T, U, a.tr, 0, e, E);
        trace4<D::NQ, NV, D::oP, D::oEl, D::oEr>(T, V, a.tr, 4, e, E);
        trace4<D::NQ, NV, D::oP, D::oE1l, D::oE1r>(T, V, a.tr, 8, e, E);
        if (a.cheb_global) {
            const double* beta = a.beta + (long)a.cls_b[e] * D::NH * D::NH;
            double G[D::NH * D::NH];
            double th, de, s1;
            n_e = cheb_setup<D>(T, V, beta, a.cheb_L, G, th, de, s1);
            if (n_e == 0) {
                atomicMin(&a.err->mfail, e + 1);
                n_e = 1;
            }
        }
    }
    if (a.cheb_global) {
        // wave maximum, then one atomic per wave
        for (int off = 32; off > 0; off >>= 1) n_e = max(n_e, __shfl_xor(n_e, off));
        if ((threadIdx.x & 63) == 0) atomicMax(&a.cnt[a.slot], n_e);
    }
}

// ---------------------------------------------------------------------------
// interior face slots of one element side. plus = 1: the element is the "+"
// element of the face (its B or L side), else the "-" element (T or R side).
// Own traces (du, v, dv) and the neighbour's facing traces are given.
template <int NQ>
__device__ __forceinline__ void interior_slots(const KArgs& a, bool plus, int fidx, bool hz,
                                               const double (&odu)[NQ], const double (&ov)[NQ],
                                               const double (&odv)[NQ], const double (&ndu)[NQ],
                                               const double (&nv)[NQ], const double (&ndv)[NQ],
                                               double (&A)[NQ], double (&Dd)[NQ], double (&Ff)[NQ]) {
    const double th = a.theta, tau = a.tau, lu = a.lu, lv = a.lv, two_h = a.two_h;
    const int st = a.fstride;
    const FaceArr& fc2 = a.f[hz ? F_HC2 : F_VC2];
    const FaceArr& faf = a.f[hz ? F_HAF : F_VAF];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        // "-" / "+" traces of this face point
        const double vm = plus ? nv[i] : ov[i], vp = plus ? ov[i] : nv[i];
        const double dum = plus ? ndu[i] : odu[i], dup = plus ? odu[i] : ndu[i];
        const double dvm = plus ? ndv[i] : odv[i], dvp = plus ? odv[i] : ndv[i];
        if (a.one_sided) {
            const double c2m = fv(a.f[hz ? F_HC2M : F_VC2M], i, fidx, st), c2p = fv(fc2, i, fidx, st);
            const double afm = fv(a.f[hz ? F_HAFM : F_VAFM], i, fidx, st), afp = fv(faf, i, fidx, st);
            const double K = fv(a.f[hz ? F_HJC : F_VJC], i, fidx, st);
            const double J = two_h * (c2m * dum - c2p * dup + afm * dvm - afp * dvp) + K;
            const double G = two_h * ((1 - th) * (c2m * dum + afm * dvm) + th * (c2p * dup + afp * dvp));
            const double pen = (0.5 * (c2m + c2p)) * lu + (0.5 * (afm + afp)) * lv;
            const double s = vm - vp;
            if (plus) {
                const double dvsp = th * s - tau * J;
                A[i] = pen * s - G - (1 - th) * K;
                Dd[i] = -afp * dvsp;
                Ff[i] = c2p * dvsp;
            } else {
                const double dvsm = (1 - th) * (vp - vm) - tau * J;
                A[i] = G - pen * s - th * K;
                Dd[i] = afm * dvsm;
                Ff[i] = -c2m * dvsm;
            }
        } else {
            const double c2 = fv(fc2, i, fidx, st), af = fv(faf, i, fidx, st);
            const double jmp = vm - vp;
            const double gst = two_h * (c2 * (th * dup + (1 - th) * dum) + af * (th * dvp + (1 - th) * dvm));
            const double tj = tau * (c2 * (two_h * (dum - dup)) + af * (two_h * (dvm - dvp)));
            if (plus) {
                A[i] = (c2 * lu + af * lv) * jmp - gst;
                Dd[i] = -af * (th * jmp - tj);
                Ff[i] = c2 * (th * jmp - tj);
            } else {
                A[i] = gst - (c2 * lu + af * lv) * jmp;
                Dd[i] = af * ((1 - th) * (vp - vm) - tj);
                Ff[i] = -c2 * ((1 - th) * (vp - vm) - tj);
            }
        }
    }
}

// boundary face slots of one element side s (src/rhs.jl:738-765); returns the
// smallest failing key of the side (0 if none).
template <int NQ>
__device__ __forceinline__ int boundary_slots(const KArgs& a, int s, int fidx, int m,
                                              const double (&du)[NQ], const double (&vt)[NQ],
                                              const double (&dv)[NQ], double (&A)[NQ], double (&Dd)[NQ],
                                              double (&Ff)[NQ]) {
    const bool hz = s < 2;
    const int st = a.fstride;
    const double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    const double* sp = a.bcd[2 * s];
    const double* spn = a.bcd[2 * s + 1];
    const int kind = a.kind[s];
    const double two_h = a.two_h, om = a.omega, ct = a.ct, stt = a.st;
    int key = 0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const double c2f = fv(a.f[hz ? F_HC2 : F_VC2], i, fidx, st);
        const double af = fv(a.f[hz ? F_HAF : F_VAF], i, fidx, st);
        const double cf = fv(a.f[hz ? F_HCF : F_VCF], i, fidx, st);
        const double bf = fv(a.f[hz ? F_HBF : F_VBF], i, fidx, st);
        const double v_fp = vt[i];
        const double gun = sgn * (two_h * du[i]);
        const double gvn = sgn * (two_h * dv[i]);
        const double spv = sp[i + NQ * m], spnv = spn[i + NQ * m];
        const double gu_ex = sgn * (ct * spnv - spnv);
        const double gv_ex = sgn * (-om * stt * spnv);
        const double v_ex = -om * stt * spv;
        double F1, dvv;
        if (kind == 3) {                                   // wall (src/rhs.jl:91-96)
            const double g2 = c2f * gu_ex + af * gv_ex;
            const double rho_w = c2f * gun + af * gvn - g2;
            F1 = g2;
            dvv = -(a.kap[s] / cf) * rho_w;
        } else {
            const double kb = a.kb[s];
            const int check = 1.0 - kb * bf * v_fp <= 0.0 ? 1 : (1.0 - kb * bf * v_ex <= 0.0 ? 2 : 0);
            if (check) {
                const int k2 = 2 * (s * NQ * a.N + i + m * NQ) + check;
                key = key == 0 ? k2 : min(key, k2);
                A[i] = 0.0; Dd[i] = 0.0; Ff[i] = 0.0;
                continue;
            }
            const double g = a.gam[s], et = a.eta[s], sg = a.sig[s];   // src/rhs.jl:74-85
            const double rho_num = et * (c2f * gun + af * gvn) + g * cf * sqrt(1.0 - kb * bf * v_fp) * v_fp;
            const double rho_ex = et * (c2f * gu_ex + af * gv_ex) + g * cf * sqrt(1.0 - kb * bf * v_ex) * v_ex;
            const double rho_net = rho_num - rho_ex;
            const double sqf = sqrt(1.0 - kb * bf * v_fp);
            F1 = c2f * gun + af * gvn - (et + sg * g) * rho_net;
            dvv = -(g - sg * et) / (cf * sqf) * rho_net;
        }
        A[i] = F1;
        Dd[i] = sgn * (af * dvv);
        Ff[i] = -sgn * (c2f * dvv);
    }
    return key;
}

// lift of one side's slots (src/rhs.jl:433-470)
template <class D>
__device__ __forceinline__ void lift_side(const double* __restrict__ T, int s, double two_h, double four_hh,
                                          const double (&A)[D::NQ], const double (&Dd)[D::NQ],
                                          const double (&Ff)[D::NQ], double (&Rv)[D::NV * D::NV],
                                          double (&Fu)[D::NU * D::NU]) {
    constexpr int NQ = D::NQ, NU = D::NU, NV = D::NV;
    const int oE = (s == 0 || s == 2) ? D::oEl : D::oEr;
    const int oE1 = (s == 0 || s == 2) ? D::oE1l : D::oE1r;
    double YA[NV], YD[NV], YF[NU];
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        double sA = 0.0, sD = 0.0;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            sA = fma(T[D::oP + i + NQ * b], T[D::oW + i] * A[i], sA);
            sD = fma(T[D::oP + i + NQ * b], T[D::oW + i] * Dd[i], sD);
        }
        YA[b] = sA; YD[b] = sD;
    }
#pragma unroll
    for (int b = 0; b < NU; ++b) {
        double sF = 0.0;
#pragma unroll
        for (int i = 0; i < NQ; ++i) sF = fma(T[D::oP + i + NQ * b], T[D::oW + i] * Ff[i], sF);
        YF[b] = sF;
    }
    if (s < 2) {
#pragma unroll
        for (int b = 0; b < NV; ++b)
#pragma unroll
            for (int aa = 0; aa < NV; ++aa) {
                double r = Rv[aa + NV * b];
                r = r + (two_h * T[oE + aa]) * YA[b];
                r = r + (four_hh * T[oE1 + aa]) * YD[b];
                Rv[aa + NV * b] = r;
            }
#pragma unroll
        for (int b = 0; b < NU; ++b)
#pragma unroll
            for (int aa = 0; aa < NU; ++aa) Fu[aa + NU * b] = Fu[aa + NU * b] + (1.0 * T[oE1 + aa]) * YF[b];
    } else {
#pragma unroll
        for (int b = 0; b < NV; ++b)
#pragma unroll
            for (int aa = 0; aa < NV; ++aa) {
                double r = Rv[aa + NV * b];
                r = r + (two_h * YA[aa]) * T[oE + b];
                r = r + (four_hh * YD[aa]) * T[oE1 + b];
                Rv[aa + NV * b] = r;
            }
#pragma unroll
        for (int b = 0; b < NU; ++b)
#pragma unroll
            for (int aa = 0; aa < NU; ++aa) Fu[aa + NU * b] = Fu[aa + NU * b] + (1.0 * YF[aa]) * T[oE1 + b];
    }
}

template <int NQ>
__device__ __forceinline__ void load_tr(const double* __restrict__ tr, int arr, int e, int E, double (&o)[NQ]) {
#pragma unroll
    for (int i = 0; i < NQ; ++i) o[i] = tr[(long)(arr * NQ + i) * E + e];
}

__device__ __forceinline__ bool skip_stage(const KArgs& a) {
    const int fl = a.err->failed;
    return fl != 0 && fl < a.serial;
}

// ---------------------------------------------------------------------------
// Volume terms (src/rhs.jl:317-333) scaled by -4/h^2 plus the source
// (src/rhs.jl:340-350) into R, and VB = (c^2 grad phi_u, grad v_h) into B.
// The point values are formed one point row p at a time and contracted back
// immediately (rank-one updates), so only the modal arrays stay live.
template <int NU, int NV, bool UNIC>
__global__ __launch_bounds__(kBlock) void k_volume(const KArgs a, double* __restrict__ R,
                                                   double* __restrict__ B) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ;
    if (skip_stage(a)) return;
    const int E = a.E;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    const double* __restrict__ T = a.tab;
    const long cc = UNIC ? 0 : (long)a.cls_ca[e];
    const double* __restrict__ Cw = a.Cw + cc * NQ * NQ;
    const double* __restrict__ Aw = a.Aw + cc * NQ * NQ;
    double Rv[NV * NV];
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) Rv[m] = 0.0;
    // (c^2 grad phi_v, grad u_h)
    {
        double U[NU * NU];
#pragma unroll
        for (int m = 0; m < NU * NU; ++m) U[m] = a.yu[(long)m * E + e];
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t1[NU], t2[NU];
#pragma unroll
            for (int b = 0; b < NU; ++b) {
                double s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int aa = 0; aa < NU; ++aa) {
                    s1 = fma(T[D::oS + p + NQ * aa], U[aa + NU * b], s1);
                    s2 = fma(T[D::oP + p + NQ * aa], U[aa + NU * b], s2);
                }
                t1[b] = s1; t2[b] = s2;
            }
            double z1[NV], z2[NV];
#pragma unroll
            for (int b = 0; b < NV; ++b) { z1[b] = 0.0; z2[b] = 0.0; }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double g1 = 0.0, g2 = 0.0;
#pragma unroll
                for (int b = 0; b < NU; ++b) {
                    g1 = fma(t1[b], T[D::oP + q + NQ * b], g1);
                    g2 = fma(t2[b], T[D::oS + q + NQ * b], g2);
                }
                const double cw = Cw[p + NQ * q];
                const double h1 = cw * g1, h2 = cw * g2;
#pragma unroll
                for (int b = 0; b < NV; ++b) {
                    z1[b] = fma(h1, T[D::oP + q + NQ * b], z1[b]);
                    z2[b] = fma(h2, T[D::oS + q + NQ * b], z2[b]);
                }
            }
#pragma unroll
            for (int b = 0; b < NV; ++b)
#pragma unroll
                for (int aa = 0; aa < NV; ++aa)
                    Rv[aa + NV * b] = fma(T[D::oS + p + NQ * aa], z1[b], fma(T[D::oP + p + NQ * aa], z2[b], Rv[aa + NV * b]));
        }
    }
    // (alpha grad phi_v, grad v_h) into Rv and (c^2 grad phi_u, grad v_h) into VB
    {
        double V[NV * NV], VB[NU * NU];
#pragma unroll
        for (int m = 0; m < NV * NV; ++m) V[m] = a.yv[(long)m * E + e];
#pragma unroll
        for (int m = 0; m < NU * NU; ++m) VB[m] = 0.0;
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            double t1[NV], t2[NV];
#pragma unroll
            for (int b = 0; b < NV; ++b) {
                double s1 = 0.0, s2 = 0.0;
#pragma unroll
                for (int aa = 0; aa < NV; ++aa) {
                    s1 = fma(T[D::oS + p + NQ * aa], V[aa + NV * b], s1);
                    s2 = fma(T[D::oP + p + NQ * aa], V[aa + NV * b], s2);
                }
                t1[b] = s1; t2[b] = s2;
            }
            double z1[NV], z2[NV], y1[NU], y2[NU];
#pragma unroll
            for (int b = 0; b < NV; ++b) { z1[b] = 0.0; z2[b] = 0.0; }
#pragma unroll
            for (int b = 0; b < NU; ++b) { y1[b] = 0.0; y2[b] = 0.0; }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double g1 = 0.0, g2 = 0.0;
#pragma unroll
                for (int b = 0; b < NV; ++b) {
                    g1 = fma(t1[b], T[D::oP + q + NQ * b], g1);
                    g2 = fma(t2[b], T[D::oS + q + NQ * b], g2);
                }
                const double aw = Aw[p + NQ * q], cw = Cw[p + NQ * q];
                const double h1 = aw * g1, h2 = aw * g2, c1 = cw * g1, c2 = cw * g2;
#pragma unroll
                for (int b = 0; b < NV; ++b) {
                    z1[b] = fma(h1, T[D::oP + q + NQ * b], z1[b]);
                    z2[b] = fma(h2, T[D::oS + q + NQ * b], z2[b]);
                }
#pragma unroll
                for (int b = 0; b < NU; ++b) {
                    y1[b] = fma(c1, T[D::oP + q + NQ * b], y1[b]);
                    y2[b] = fma(c2, T[D::oS + q + NQ * b], y2[b]);
                }
            }
#pragma unroll
            for (int b = 0; b < NV; ++b)
#pragma unroll
                for (int aa = 0; aa < NV; ++aa)
                    Rv[aa + NV * b] = fma(T[D::oS + p + NQ * aa], z1[b], fma(T[D::oP + p + NQ * aa], z2[b], Rv[aa + NV * b]));
#pragma unroll
            for (int b = 0; b < NU; ++b)
#pragma unroll
                for (int aa = 0; aa < NU; ++aa)
                    VB[aa + NU * b] = fma(T[D::oS + p + NQ * aa], y1[b], fma(T[D::oP + p + NQ * aa], y2[b], VB[aa + NU * b]));
        }
#pragma unroll
        for (int m = 0; m < NU * NU; ++m) B[(long)m * E + e] = VB[m];
    }
    // scale and source, in the order of the reference's sum
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) {
        double r = (-a.four_hh) * Rv[m];
        for (int t = 0; t < a.nproj; ++t) r = r + a.pc[t] * a.proj[t][(long)m * E + e];
        R[(long)m * E + e] = r;
    }
}

// ---------------------------------------------------------------------------
// Face slots of the element's four sides from its own and its neighbours'
// traces, lifted into R and Fu (src/rhs.jl:421-473, 589-776); then the
// u-equation with the packed inverse (src/rhs.jl:361-382) and the u part of
// the RK4 stage update (src/solver.jl:300-321).
template <int NU, int NV, bool UNIC>
__global__ __launch_bounds__(kBlock) void k_faces(const KArgs a, double* __restrict__ R,
                                                  const double* __restrict__ B) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ;
    if (skip_stage(a)) return;
    const int E = a.E, N = a.N;
    const int mfail_prev = a.cheb_global ? a.err->mfail : 0x7fffffff;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    const double* __restrict__ T = a.tab;
    const int j = e % N, k = e / N;
    double Rv[NV * NV], Fu[NU * NU];
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) Rv[m] = R[(long)m * E + e];
#pragma unroll
    for (int m = 0; m < NU * NU; ++m) Fu[m] = 0.0;
    int bkey = 0;
#pragma unroll 1
    for (int s = 0; s < 4; ++s) {
        const bool hz = s < 2;
        const bool plus = (s == 0 || s == 2);
        double odu[NQ], ov[NQ], odv[NQ];
        load_tr<NQ>(a.tr, 0 + s, e, E, odu);
        load_tr<NQ>(a.tr, 4 + s, e, E, ov);
        load_tr<NQ>(a.tr, 8 + s, e, E, odv);
        int fidx, nb, m;
        bool bnd;
        if (s == 0)      { fidx = j + (N + 1) * k;     nb = e - 1; bnd = j == 0;     m = k; }
        else if (s == 1) { fidx = j + 1 + (N + 1) * k; nb = e + 1; bnd = j == N - 1; m = k; }
        else if (s == 2) { fidx = j + N * k;           nb = e - N; bnd = k == 0;     m = j; }
        else             { fidx = j + N * (k + 1);     nb = e + N; bnd = k == N - 1; m = j; }
        double A[NQ], Dd[NQ], Ff[NQ];
        if (bnd) {
            const int k2 = boundary_slots<NQ>(a, s, fidx, m, odu, ov, odv, A, Dd, Ff);
            if (k2) bkey = bkey == 0 ? k2 : min(bkey, k2);
        } else {
            const int ns = s ^ 1;   // the neighbour's facing side: B <-> T, L <-> R
            double ndu[NQ], nvv[NQ], ndv[NQ];
            load_tr<NQ>(a.tr, 0 + ns, nb, E, ndu);
            load_tr<NQ>(a.tr, 4 + ns, nb, E, nvv);
            load_tr<NQ>(a.tr, 8 + ns, nb, E, ndv);
            interior_slots<NQ>(a, plus, fidx, hz, odu, ov, odv, ndu, nvv, ndv, A, Dd, Ff);
        }
        lift_side<D>(T, s, a.two_h, a.four_hh, A, Dd, Ff, Rv, Fu);
    }
    if (bkey) {
        atomicMin(&a.err->bkey, bkey);
        a.err->failed = a.serial;
    }
    if (mfail_prev != 0x7fffffff) a.err->failed = a.serial;
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) R[(long)m * E + e] = Rv[m];

    // u-equation u_t = A_e^-1 b (src/rhs.jl:361-382). The inverse of the class is
    // stored unpacked; it is symmetric, so row i is the contiguous column i and
    // each output is finished (and written) before the next one starts.
    constexpr int n = NU * NU;
    const long cc = UNIC ? 0 : (long)a.cls_ca[e];
    const double* __restrict__ Ai = a.Ainv + cc * (n * n);
    double b[n];
#pragma unroll
    for (int i = 0; i < n; ++i) b[i] = B[(long)i * E + e] - Fu[i];
    b[0] += a.yv[e];
    // RK4 stage update, u part (src/solver.jl:300-321)
    const double c = a.cstage;
    bool fin = true;
#pragma unroll
    for (int mm = 0; mm < n; ++mm) {
        double kv = 0.0;
#pragma unroll
        for (int jj = 0; jj < n; ++jj) kv = fma(Ai[jj + n * mm], b[jj], kv);
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
// The mass system M(v_h) X = R by Chebyshev iteration (src/rhs.jl:886-939)
// and the v part of the RK4 stage update.
template <int NU, int NV>
__global__ __launch_bounds__(kBlock) void k_mass(const KArgs a, const double* __restrict__ R) {
    using D = Dims<NU, NV>;
    constexpr int NH = D::NH;
    if (skip_stage(a)) return;
    const int E = a.E;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.cnt[a.slot ^ 1] = 1;   // the next stage's count
    const int e = blockIdx.x * kBlock + threadIdx.x;
    int ncheb = 0;
    if (e < E) {
        const double* __restrict__ T = a.tab;
        double G[NH * NH];
        double th = 1.0, de = 1.0, s1 = 1.0;
        int n_e;
        {
            double V[NV * NV];
#pragma unroll
            for (int m = 0; m < NV * NV; ++m) V[m] = a.yv[(long)m * E + e];
            const double* beta = a.beta + (long)a.cls_b[e] * NH * NH;
            n_e = cheb_setup<D>(T, V, beta, a.cheb_L, G, th, de, s1);
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
#pragma unroll
            for (int m = 0; m < NH * NH; ++m) G[m] *= T[D::oWh + m];
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
                // r <- (r - d) + K d,  K d = P_h^T (G o (P_h d P_h^T)) P_h
#pragma unroll
                for (int m = 0; m < NV * NV; ++m) r[m] = r[m] - d[m];
#pragma unroll
                for (int p = 0; p < NH; ++p) {
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
                        y *= G[p + NH * q];
#pragma unroll
                        for (int b = 0; b < NV; ++b) zp[b] = fma(y, T[D::oPh + q + NH * b], zp[b]);
                    }
#pragma unroll
                    for (int b = 0; b < NV; ++b)
#pragma unroll
                        for (int aa = 0; aa < NV; ++aa)
                            r[aa + NV * b] = fma(T[D::oPh + p + NH * aa], zp[b], r[aa + NV * b]);
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
    if ((threadIdx.x & 63) == 0 && ncheb > 0) atomicMax(&a.err->maxcheb, ncheb);
}

}  // namespace wedg
