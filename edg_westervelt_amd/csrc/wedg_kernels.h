// Device kernels of the EDG Westervelt right-hand side and RK4 stages (gfx950, FP64).
//
// One lane owns one element e = j + N k: it reads the element's modal
// coefficients straight from the state layout (N, N, n, n) — consecutive lanes
// read consecutive elements, so every load and store is one coalesced wave
// access — and keeps the element's small matrices in registers. The 1D
// Legendre tables live in a read-only buffer read through the constant
// address space, so they arrive by scalar loads and feed the FMAs as SGPR
// operands.
//
// The tables are values (or derivatives) of Legendre polynomials at Gauss
// points, which are mirrored about the centre: T[n-1-p][a] = (-1)^a T[p][a]
// for values and -(-1)^a T[p][a] for derivatives. Every contraction is split
// into its even and odd modes (rows_sym / cols_sym), which halves its FMAs; the
// results differ from the plain sums by rounding only.
//
// One kernel per RK stage, k_stage: each element forms the slots of its
// physical-boundary sides with their weight checks (src/rhs.jl:65-128,
// 702-776) from its stage state, its own traces and its
// neighbours' facing ones from the state (reference _traces!/_trace4!,
// src/rhs.jl:384-410, 573-580), its face slots (interior fluxes
// src/rhs.jl:620-700), the volume terms (src/rhs.jl:317-333) and source
// (src/rhs.jl:340-350), the lifts (src/rhs.jl:421-473), the u-equation
// (src/rhs.jl:361-382), the Chebyshev mass solve (src/rhs.jl:889-939) and the
// RK4 stage update (src/solver.jl:300-321), and then, from the state it has
// just written, what the next stage needs of all elements before it can start:
// the a-priori Chebyshev count of every element reduced per wave (the global
// policy takes the maximum, src/rhs.jl:889-918). k_prep does the same for the
// first stage of a call.
// Every face slot is computed by the element that owns it, so no slot has two
// writers and no floating-point atomics exist; results do not depend on the
// launch shape.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace wedg {

constexpr int kWave = 64;             // lanes per block: one wave
constexpr int kMassLdsDoubles = 80;   // Chebyshev weights per lane kept in LDS: 4 waves fill 160 KB

// Coefficient layouts the element kernel is specialised for.
enum {
    MODE_GENERAL = 0,     // per-element classes, stored inverse of the u-stiffness
    MODE_SEPARABLE = 1,   // per-element classes, c^2 constant on each element: spectral u-solve
    MODE_UNIFORM = 2,     // one class, one beta record, every face array one value; spectral u-solve
    MODE_GENERAL_CG = 3,  // as MODE_GENERAL, mass systems by conjugate gradients (mass_solver = :cg)
};

template <int NU_, int NV_, bool BT_ = false>
struct Dims {
    static constexpr bool BT = BT_;   // a batched launch (BatchArgs): the block's instance selects its KArgs
    static constexpr int NU = NU_, NV = NV_, NQ = NU_;
    static constexpr int QU = NU_ - 1, QV = NV_ - 1;
    static constexpr int NH0 = (3 * QV + 1) / 2 + 1;                 // ceil(3 q_v / 2) + 1
    static constexpr int NH = NH0 > NU_ ? NH0 : NU_;                  // src/solver.jl:483
    static constexpr int NP = NH / 2;                                 // mirrored pairs of high-order rows
    static constexpr int PL0 = kMassLdsDoubles / (2 * NH);
    static constexpr int PL = PL0 < NP ? PL0 : NP;                    // pairs of G rows in LDS
    static constexpr int PR = NP - PL;                                // pairs of G rows in registers
    static constexpr int NYL = 4 * (2 * NV_ + NU_);                   // lift vectors of the four sides
    static constexpr int NLDS0 = 2 * PL > 0 ? 2 * PL * NH : 1;
    static constexpr int NLDS = NLDS0 > NYL ? NLDS0 : NYL;            // LDS doubles per lane
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
    static constexpr int oQ = oWh + NH * NH;        // eigenvectors of the 1D stiffness (NU x NU)
    static constexpr int oTEl = oQ + NU * NU;       // (2/h) Ebl, (2/h) Ebr, (4/h^2) E1bl, (4/h^2) E1br (NU each)
    static constexpr int oTEr = oTEl + NU;
    static constexpr int oFE1l = oTEr + NU;
    static constexpr int oFE1r = oFE1l + NU;
    static constexpr int nTab = oFE1r + NU;
};

struct FaceArr {
    const double* p;   // [i][face] (face index fastest), or one value
    int uni;           // 1: the array is one constant
    int pad;
};

// Failure record of the device (reset by the host before each public call).
struct DevErr {
    // The first failure in serial order (all ones = none): the stage's serial,
    // then the kind (0: a boundary weight, before 1: the mass matrix), then the
    // key (the _boundary_key, or the element + 1), packed so that atomicMin
    // keeps the failure the reference reports.
    unsigned long long fkey;
    int nonfinite;     // first RK4 step with a non-finite entry (INT_MAX = none)
    int cg_unconverged;   // element systems conjugate gradients left above the tolerance
    int nf_serial;     // serial of the stage that made the first non-finite state (INT_MAX = none)
    int pad[3];
};

enum { F_HC2 = 0, F_HAF, F_HCF, F_HBF, F_VC2, F_VAF, F_VCF, F_VBF,
       F_HC2M, F_HAFM, F_VC2M, F_VAFM, F_HJC, F_VJC, F_COUNT };

struct KArgs {
    int N, E, stage, cheb_global;
    int serial, slot, one_sided, step;
    int kind[4];
    int nproj, mass_cg;             // mass_cg: conjugate gradients instead of Chebyshev iteration
    double two_h, four_hh, theta, tau, lu, lv, omega, cheb_L;
    double mass_tol;
    double ct, st, srct, cstage;
    double nct, nst;                // time factors of the next stage (its boundary data)
    int has_next, pad2;             // 1: prepare the next stage from this stage's output
    double pc[5];                   // coefficients of the projections
    double gam[4], eta[4], sig[4], kb[4], kap[4];
    const double* tab;
    const int* cls_ca;              // class of (Cw, Aw, u-solve record) per element
    const double* Cw;               // per class (NQ*NQ), iy + NQ ix
    const double* Aw;
    const double* Ainv;             // per class: the inverse (NU^2 x NU^2, symmetric) or, for the
                                    // spectral u-solve, its eigenvalue factors D (NU x NU)
    const int* cls_b;               // class of beta_hq per element
    const double* beta;             // per class (NH*NH), p + NH q
    FaceArr f[F_COUNT];
    double fu[F_COUNT];             // the value of each face array that is one constant
    int fstride;                    // N (N + 1): stride of the point index of face arrays
    int pad1;
    const double* bcd[8];           // boundary data (NQ, N): sp, spn for bot, top, lft, rgt
    const double* proj[5];          // element-fastest [m][E]
    const double* yu; const double* yv;     // stage state Y_s
    const double* y0u; const double* y0v;   // base state y of the step
    double* accu; double* accv;             // RK4 accumulator
    double* outu; double* outv;             // Y_{s+1}, y (last stage) or (u_t, v_t)
    double* bsl;                            // boundary slots A, D, F [side][i][3][N]
    DevErr* err;
    int* cnt;                               // Chebyshev count of every wave [2][waves], by stage parity
    unsigned* wstat;                        // per wave: sum [waves] and max [waves] of the counts it ran
    double* recu; double* recv;             // when not null: the stage's derivative k = F(Y_s) is kept here
    const double* tab2;                     // the split tables of the two-lane kernels (wedg_stage2.h) or null
};

// ---------------------------------------------------------------------------
// Table access.
#if defined(__HIP_DEVICE_COMPILE__)
#define WEDG_CONST_AS __attribute__((address_space(4)))
#else
#define WEDG_CONST_AS
#endif
typedef const WEDG_CONST_AS double cdbl;

// The table pointer in the constant address space, opaque to the optimizer:
// loads through it are scalar and are issued where they are used instead of
// being hoisted out of a loop and kept alive for the whole kernel.
__device__ __forceinline__ cdbl* ctab(const double* p) {
    cdbl* q = (cdbl*)p;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(q));
#endif
    return q;
}


// A batched launch advances the instances of one (n_u, n_v, mode) group at
// once: each instance owns a contiguous range of blocks and its own KArgs.
struct BatchArgs {
    const KArgs* args;    // [instances] this stage's arguments of every instance
    const int* binst;     // [blocks] the instance of each block
    const int* bfirst;    // [instances] the first block of each instance
};

// The kernel's own arguments read through the kernarg segment pointer in the
// constant address space, opaque to the optimizer like ctab(): fields are
// loaded (scalar) where they are used instead of being kept in SGPRs across
// the whole kernel. A plain kernel takes one KArgs by value, at offset 0; a
// batched one (BT) a BatchArgs, whose block-to-instance lookup is a pair of
// scalar loads from read-only memory that the compiler forms once.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const WEDG_CONST_AS KArgs CKArgs;
typedef const WEDG_CONST_AS BatchArgs CBatchArgs;
typedef const WEDG_CONST_AS int cint;
__device__ __forceinline__ CBatchArgs* bargs() { return (CBatchArgs*)__builtin_amdgcn_kernarg_segment_ptr(); }
__device__ __forceinline__ int block_instance() { return ((cint*)bargs()->binst)[blockIdx.x]; }
template <bool BT = false>
__device__ __forceinline__ CKArgs* kargs() {
    CKArgs* p;
    if constexpr (BT) p = (CKArgs*)bargs()->args + block_instance();
    else p = (CKArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return p;
}
#else
typedef const KArgs CKArgs;
typedef const BatchArgs CBatchArgs;
typedef const int cint;
extern const KArgs* wedg_emu_kargs;        // set by a host build that runs the kernels on the CPU
extern const BatchArgs* wedg_emu_bargs;
__device__ __forceinline__ CBatchArgs* bargs() { return wedg_emu_bargs; }
__device__ __forceinline__ int block_instance() { return bargs()->binst[blockIdx.x]; }
template <bool BT = false>
__device__ __forceinline__ CKArgs* kargs() {
    if constexpr (BT) return bargs()->args + block_instance();
    else return wedg_emu_kargs;
}
#endif

// A per-lane pointer, opaque to the optimizer: loads through it are not
// hoisted above this point (and kept in registers until their use).
__device__ __forceinline__ const double* vfresh(const double* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(p));
#endif
    return p;
}

// Element-fastest arrays x[m][E] of doubles through buffer instructions: the
// lane's byte offset 8 e sits in one VGPR for every mode m and the mode's
// byte offset 8 m E in an SGPR, so an access needs no 64-bit address
// arithmetic and no address registers of its own (arrays below 2 GiB).
#if defined(__HIP_DEVICE_COMPILE__)
typedef unsigned int wedg_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t ersrc(const double* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), (short)0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ double eld(__amdgpu_buffer_rsrc_t r, int e, int m, int E) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, e * 8, m * E * 8, 0));
}
__device__ __forceinline__ void est(__amdgpu_buffer_rsrc_t r, int e, int m, int E, double x) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(wedg_u32x2, x), r, e * 8, m * E * 8, 0);
}
template <int n>
__device__ __forceinline__ void eload(const double* p, int E, int e, double (&x)[n]) {
    const __amdgpu_buffer_rsrc_t r = ersrc(p);
#pragma unroll
    for (int m = 0; m < n; ++m) x[m] = eld(r, e, m, E);
}
template <int n>
__device__ __forceinline__ void estore(double* p, int E, int e, const double (&x)[n]) {
    const __amdgpu_buffer_rsrc_t r = ersrc(p);
#pragma unroll
    for (int m = 0; m < n; ++m) est(r, e, m, E, x[m]);
}
#else
template <int n>
__device__ __forceinline__ void eload(const double* p, int E, int e, double (&x)[n]) {
    for (int m = 0; m < n; ++m) x[m] = p[(long)m * E + e];
}
template <int n>
__device__ __forceinline__ void estore(double* p, int E, int e, const double (&x)[n]) {
    for (int m = 0; m < n; ++m) p[(long)m * E + e] = x[m];
}
#endif

// value of face array x at flat index fi = i * (N (N + 1)) + face
template <bool UNI, class KA>
__device__ __forceinline__ double fval(const KA& a, int x, long fi) {
    if (UNI) return a.fu[x];
    return a.f[x].uni ? a.fu[x] : a.f[x].p[fi];
}

// A phase boundary inside a kernel: neither the optimizer nor the scheduler
// moves loads or arithmetic of a later phase into an earlier one, which
// would hold the later phase's operands in registers through the earlier one.
__device__ __forceinline__ void phase_barrier() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
#endif
}

// a failure of an earlier stage ends the run
template <class KA>
__device__ __forceinline__ bool skip_stage(const KA& a) {
    return (a.err->fkey >> 33) < (unsigned long long)(unsigned)a.serial;
}

// First element of the block. Blocks are dealt to the 8 XCDs in turn; this
// gives each XCD a contiguous band of elements, so the neighbours across k
// live in the same L2 except at the band edges. In a batched launch, the
// block's place in its instance's range.
template <bool BT = false>
__device__ __forceinline__ int block_first_element() {
    if constexpr (BT) {
        return ((int)blockIdx.x - ((cint*)bargs()->bfirst)[block_instance()]) * kWave;
    } else {
        const int b = blockIdx.x, nb = gridDim.x;
        const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
        return lb * kWave;
    }
}

// ---------------------------------------------------------------------------
// Mirrored contractions. T is column-major with NR rows (points) and at least
// n columns (modes); ODD marks a derivative table.

// rows p and NR-1-p (p < NR/2) of sum_a T[r, a] x(a)
template <int NR, int n, bool ODD, class X>
__device__ __forceinline__ void row_pair(cdbl* T, int p, X x, double& yp, double& yq) {
    double ev = 0.0, od = 0.0;
#pragma unroll
    for (int a = 0; a < n; a += 2) ev = fma(T[p + NR * a], x(a), ev);
#pragma unroll
    for (int a = 1; a < n; a += 2) od = fma(T[p + NR * a], x(a), od);
    yp = ev + od;
    yq = ODD ? od - ev : ev - od;
}

// the middle row (NR odd): only the modes of one parity are non-zero there
template <int NR, int n, bool ODD, class X>
__device__ __forceinline__ double row_mid(cdbl* T, X x) {
    constexpr int m = NR / 2;
    double s = 0.0;
#pragma unroll
    for (int a = ODD ? 1 : 0; a < n; a += 2) s = fma(T[m + NR * a], x(a), s);
    return s;
}

// y[r] = sum_a T[r, a] x(a) for every row r
template <int NR, int n, bool ODD, class X>
__device__ __forceinline__ void rows_sym(cdbl* T, X x, double (&y)[NR]) {
#pragma unroll
    for (int p = 0; p < NR / 2; ++p) row_pair<NR, n, ODD>(T, p, x, y[p], y[NR - 1 - p]);
    if (NR & 1) y[NR / 2] = row_mid<NR, n, ODD>(T, x);
}

// z[a] = sum_r T[r, a] y[r] for the first n columns
template <int NR, int n, bool ODD>
__device__ __forceinline__ void cols_sym(cdbl* T, const double (&y)[NR], double (&z)[n]) {
    constexpr int H = NR / 2;
    double s[H > 0 ? H : 1], d[H > 0 ? H : 1];
#pragma unroll
    for (int p = 0; p < H; ++p) {
        s[p] = y[p] + y[NR - 1 - p];
        d[p] = y[p] - y[NR - 1 - p];
    }
#pragma unroll
    for (int a = 0; a < n; ++a) {
        const bool same = ((a & 1) == 0) != ODD;   // T[NR-1-p, a] == T[p, a]
        double acc = 0.0;
#pragma unroll
        for (int p = 0; p < H; ++p) acc = fma(T[p + NR * a], same ? s[p] : d[p], acc);
        if ((NR & 1) && same) acc = fma(T[H + NR * a], y[H], acc);
        z[a] = acc;
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

// Rows p and NH-1-p of v_q = P_h V P_h^T (the values of v_h at the high-order
// points); the middle row when p == NH/2 and NH is odd (then yq is unused).
template <class D>
__device__ __forceinline__ void vq_rows(cdbl* Ph, int p, bool mid, const double (&V)[D::NV * D::NV],
                                        double (&yp)[D::NH], double (&yq)[D::NH]) {
    constexpr int NV = D::NV, NH = D::NH;
    double tp[NV], tq[NV];
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        if (mid) {
            tp[b] = row_mid<NH, NV, false>(Ph, [&](int a) { return V[a + NV * b]; });
            tq[b] = 0.0;
        } else {
            row_pair<NH, NV, false>(Ph, p, [&](int a) { return V[a + NV * b]; }, tp[b], tq[b]);
        }
    }
    rows_sym<NH, NV, false>(Ph, [&](int b) { return tp[b]; }, yp);
    if (!mid) rows_sym<NH, NV, false>(Ph, [&](int b) { return tq[b]; }, yq);
}

// ---------------------------------------------------------------------------
// Slot values (A, D, F) of one face point of an interior face, for the element
// on the "+" side (plus) or the "-" side (src/rhs.jl:620-700). vm, vp, ... are
// the traces of the "-" and "+" elements at the point; fi indexes the face
// arrays at the point.
template <bool UNI, class KA>
__device__ __forceinline__ void interior_point(const KA& a, bool plus, bool hz, long fi, double vm, double vp,
                                               double dum, double dup, double dvm, double dvp, double& A,
                                               double& Dd, double& Ff) {
    const double th = a.theta, tau = a.tau, lu = a.lu, lv = a.lv, two_h = a.two_h;
    const int fc2 = hz ? F_HC2 : F_VC2, faf = hz ? F_HAF : F_VAF;
    if (!UNI && a.one_sided) {
        const double c2m = fval<UNI>(a, hz ? F_HC2M : F_VC2M, fi), c2p = fval<UNI>(a, fc2, fi);
        const double afm = fval<UNI>(a, hz ? F_HAFM : F_VAFM, fi), afp = fval<UNI>(a, faf, fi);
        const double K = fval<UNI>(a, hz ? F_HJC : F_VJC, fi);
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
        const double c2 = fval<UNI>(a, fc2, fi), af = fval<UNI>(a, faf, fi);
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
// returns the point's failing _boundary_key (src/rhs.jl:732) or 0. HZ: s is a
// horizontal side (0 or 1, possibly different in different lanes).
// The data of boundary point i of side s: c^2, alpha, c, beta at the point and
// the boundary data (sp, spn). A side's points are all loaded before any of
// its slots is stored: the stores may alias these arrays as far as the
// compiler knows, and would otherwise hold each point's loads until the
// previous point's stores.
struct BPoint {
    double c2f, af, cf, bf, sp, spn;
};

template <bool UNI, bool HZ, class KA>
__device__ __forceinline__ BPoint boundary_data(const KA& a, int s, long fi, int i, int m, int nq) {
    constexpr bool hz = HZ;
    BPoint d;
    d.c2f = fval<UNI>(a, hz ? F_HC2 : F_VC2, fi);
    d.af = fval<UNI>(a, hz ? F_HAF : F_VAF, fi);
    d.cf = fval<UNI>(a, hz ? F_HCF : F_VCF, fi);
    d.bf = fval<UNI>(a, hz ? F_HBF : F_VBF, fi);
    const double* bsp = HZ ? (s == 0 ? a.bcd[0] : a.bcd[2]) : a.bcd[2 * s];
    const double* bspn = HZ ? (s == 0 ? a.bcd[1] : a.bcd[3]) : a.bcd[2 * s + 1];
    d.sp = bsp[i + nq * m];
    d.spn = bspn[i + nq * m];
    return d;
}

template <bool UNI, bool HZ, class KA>
__device__ __forceinline__ int boundary_point(const KA& a, int s, const BPoint& d, int i, int m, int nq, double du,
                                              double vt, double dv, double ct, double stt, double& A, double& Dd,
                                              double& Ff) {
    // the side's parameters: with HZ a choice between the two horizontal sides'
    // (uniform values), so that no lane-indexed load is needed
    auto sv = [&](const auto& arr) { return HZ ? (s == 0 ? arr[0] : arr[1]) : arr[s]; };
    const double sgn = (s == 0 || s == 2) ? -1.0 : 1.0;
    const double c2f = d.c2f, af = d.af, cf = d.cf, bf = d.bf;
    const double two_h = a.two_h, om = a.omega;
    const double gun = sgn * (two_h * du);
    const double gvn = sgn * (two_h * dv);
    const double spv = d.sp, spnv = d.spn;
    const double gu_ex = sgn * (ct * spnv - spnv);
    const double gv_ex = sgn * (-om * stt * spnv);
    const double v_ex = -om * stt * spv;
    double F1, dvv;
    if (sv(a.kind) == 3) {                                 // wall (src/rhs.jl:91-96)
        const double g2 = c2f * gu_ex + af * gv_ex;
        const double rho_w = c2f * gun + af * gvn - g2;
        F1 = g2;
        dvv = -(sv(a.kap) / cf) * rho_w;
    } else {
        const double kb = sv(a.kb);
        const int check = 1.0 - kb * bf * vt <= 0.0 ? 1 : (1.0 - kb * bf * v_ex <= 0.0 ? 2 : 0);
        if (check) {
            A = 0.0; Dd = 0.0; Ff = 0.0;
            return 2 * (s * nq * a.N + i + m * nq) + check;
        }
        const double g = sv(a.gam), et = sv(a.eta), sg = sv(a.sig);   // src/rhs.jl:74-85
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

// trace of the field X (n x n) on side sd of its element (0 B, 1 T, 2 L, 3 R;
// src/rhs.jl:388-410): o[i] = sum_b P[i, b] sum_a E[a] X[a, b] with E the row
// of the side's end (B, T contract the first index, L, R the second). Erow is
// the + end's row (Ebr or E1br); the - end's is (-1)^a times it for values,
// -(-1)^a for derivatives (DER).
template <class D, int n, bool DER, int sd>
__device__ __forceinline__ void side_trace(cdbl* T, int oEr, const double (&X)[n * n], double (&o)[D::NQ]) {
    constexpr int NQ = D::NQ;
    constexpr bool second = sd >= 2, plus_end = (sd & 1) != 0;
    double sv[n];
#pragma unroll
    for (int b = 0; b < n; ++b) {
        double ev = 0.0, od = 0.0;
#pragma unroll
        for (int a = 0; a < n; ++a) {
            const double x = second ? X[b + n * a] : X[a + n * b];
            if (a & 1) od = fma(T[oEr + a], x, od);
            else ev = fma(T[oEr + a], x, ev);
        }
        sv[b] = plus_end ? ev + od : (DER ? od - ev : ev - od);
    }
    rows_sym<NQ, n, false>(T + D::oP, [&](int b) { return sv[b]; }, o);
}

// side_trace of a horizontal side (B or T) whose end is chosen per lane (plus_end: T)
template <class D, int n, bool DER>
__device__ __forceinline__ void side_trace_h(cdbl* T, int oEr, const double (&X)[n * n], double (&o)[D::NQ],
                                             bool plus_end) {
    constexpr int NQ = D::NQ;
    double sv[n];
#pragma unroll
    for (int b = 0; b < n; ++b) {
        double ev = 0.0, od = 0.0;
#pragma unroll
        for (int a = 0; a < n; ++a) {
            const double x = X[a + n * b];
            if (a & 1) od = fma(T[oEr + a], x, od);
            else ev = fma(T[oEr + a], x, ev);
        }
        sv[b] = plus_end ? ev + od : (DER ? od - ev : ev - od);
    }
    rows_sym<NQ, n, false>(T + D::oP, [&](int b) { return sv[b]; }, o);
}

__device__ __forceinline__ unsigned long long fail_key(int serial, int kind, int key) {
    return ((unsigned long long)(unsigned)serial << 33) | ((unsigned long long)kind << 32) | (unsigned)key;
}

// What every element of the stage of serial `serial` must know of the others
// before the stage starts: its a-priori Chebyshev count from its input v (V)
// (with the global policy the stage takes the maximum over all elements;
// src/rhs.jl:889-918). A non-positive mass matrix is recorded with the stage's
// serial (kind 1, the element). Returns the element's count (1 when not
// counted or failing).
template <class D, bool UNI>
__device__ __forceinline__ int stage_prep(const double (&V)[D::NV * D::NV], int e, int serial, bool count) {
    constexpr int NH = D::NH;
    int n_e = 1;
    if (count) {
        CKArgs& A = *kargs<D::BT>();
        const double* betav = A.beta + (UNI ? 0 : (long)A.cls_b[e] * NH * NH);
        cdbl* Ph = ctab(A.tab) + D::oPh;
        cdbl* bu = ctab(A.beta);
        double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
        bool nan = false;
#pragma unroll
        for (int p = 0; p < (NH + 1) / 2; ++p) {
            const bool mid = (NH & 1) && p == NH / 2;
            double yp[NH], yq[NH];
            vq_rows<D>(Ph, p, mid, V, yp, yq);
#pragma unroll
            for (int q = 0; q < NH; ++q) {
                const double g = yp[q] * (2.0 * (UNI ? bu[p + NH * q] : betav[p + NH * q]));
                gmin = fmin(gmin, g);
                gmax = fmax(gmax, g);
                nan = nan || (g != g);
                if (!mid) {
                    const int pq = NH - 1 - p;
                    const double g2 = yq[q] * (2.0 * (UNI ? bu[pq + NH * q] : betav[pq + NH * q]));
                    gmin = fmin(gmin, g2);
                    gmax = fmax(gmax, g2);
                    nan = nan || (g2 != g2);
                }
            }
        }
        double th, de, s1;
        n_e = cheb_interval(gmin, gmax, nan, A.cheb_L, th, de, s1);
        if (n_e == 0) {
            atomicMin(&A.err->fkey, fail_key(serial, 1, e + 1));
            n_e = 1;
        }
    }
    return n_e;
}

// The slots A, D, F of the element's physical-boundary sides (src/rhs.jl:738-765)
// at the stage's time factors (ct, st), from its stage state (U, V), into the
// boundary slot buffer, with the weight checks of the boundary (failures kind 0,
// the _boundary_key of src/rhs.jl:732, with the stage's serial; they come
// before the stage's mass-matrix failure, kind 1).
template <class D, bool UNI>
__device__ __forceinline__ void stage_boundary(const double (&U)[D::NU * D::NU], const double (&V)[D::NV * D::NV],
                                               int e, int serial, double ct, double st) {
    constexpr int NQ = D::NQ, NU = D::NU, NV = D::NV;
    CKArgs& B = *kargs<D::BT>();
    const int N = B.N, j = e % N, k = e / N;
    if (j == 0 || j == N - 1 || k == 0 || k == N - 1) {
        int bkey = 0;
        auto side = [&](auto sc) {
            constexpr int s = decltype(sc)::value;
            const bool bnd = s == 0 ? j == 0 : s == 1 ? j == N - 1 : s == 2 ? k == 0 : k == N - 1;
            if (!bnd) return;
            CKArgs& A = *kargs<D::BT>();
            cdbl* T = ctab(A.tab);
            double du[NQ], vt[NQ], dv[NQ];
            side_trace<D, NU, true, s>(T, D::oE1r, U, du);
            side_trace<D, NV, false, s>(T, D::oEr, V, vt);
            side_trace<D, NV, true, s>(T, D::oE1r, V, dv);
            const int m = s < 2 ? k : j;
            const int fidx = s == 0 ? (N + 1) * k : s == 1 ? N + (N + 1) * k : s == 2 ? j : j + N * N;
            double* bs = A.bsl + (long)s * NQ * 3 * N + m;
            BPoint pd[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) pd[i] = boundary_data<UNI, (s < 2)>(A, s, (long)i * A.fstride + fidx, i, m, NQ);
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                double Av, Dd, Ff;
                const int k2 = boundary_point<UNI, (s < 2)>(A, s, pd[i], i, m, NQ, du[i], vt[i], dv[i], ct, st, Av, Dd, Ff);
                if (k2) bkey = bkey == 0 ? k2 : min(bkey, k2);
                bs[(long)(i * 3 + 0) * N] = Av;
                bs[(long)(i * 3 + 1) * N] = Dd;
                bs[(long)(i * 3 + 2) * N] = Ff;
            }
        };
        // The bottom (j = 0) and top (j = N - 1) sides in one pass, each lane its own:
        // in a wave of one column they are two lanes, which would otherwise run the
        // side's code one after the other. (N = 1: one element has both.)
        auto hside = [&](int s) {
            CKArgs& A = *kargs<D::BT>();
            cdbl* T = ctab(A.tab);
            const bool pe = s == 1;
            double du[NQ], vt[NQ], dv[NQ];
            side_trace_h<D, NU, true>(T, D::oE1r, U, du, pe);
            side_trace_h<D, NV, false>(T, D::oEr, V, vt, pe);
            side_trace_h<D, NV, true>(T, D::oE1r, V, dv, pe);
            const int m = k;
            const int fidx = pe ? N + (N + 1) * k : (N + 1) * k;
            double* bs = A.bsl + (long)s * NQ * 3 * N + m;
            BPoint pd[NQ];
#pragma unroll
            for (int i = 0; i < NQ; ++i) pd[i] = boundary_data<UNI, true>(A, s, (long)i * A.fstride + fidx, i, m, NQ);
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                double Av, Dd, Ff;
                const int k2 = boundary_point<UNI, true>(A, s, pd[i], i, m, NQ, du[i], vt[i], dv[i], ct, st, Av, Dd, Ff);
                if (k2) bkey = bkey == 0 ? k2 : min(bkey, k2);
                bs[(long)(i * 3 + 0) * N] = Av;
                bs[(long)(i * 3 + 1) * N] = Dd;
                bs[(long)(i * 3 + 2) * N] = Ff;
            }
        };
        if (N > 1) {
            if (j == 0 || j == N - 1) hside(j == 0 ? 0 : 1);
        } else {
            side(std::integral_constant<int, 0>());
            side(std::integral_constant<int, 1>());
        }
        side(std::integral_constant<int, 2>());
        side(std::integral_constant<int, 3>());
        if (bkey) atomicMin(&B.err->fkey, fail_key(serial, 0, bkey));
    }
}


// Writes the wave's largest count into the per-wave slot of the stage parity.
template <bool BT>
__device__ __forceinline__ void wave_count(int n_e, int slot, int E) {
    for (int off = 32; off > 0; off >>= 1) n_e = max(n_e, __shfl_xor(n_e, off));
    if ((threadIdx.x & 63) == 0) {
        const int nw = (E + kWave - 1) / kWave;
        kargs<BT>()->cnt[slot * nw + (block_first_element<BT>() / kWave)] = n_e;
    }
}

// The first stage of a call: the Chebyshev counts of its input state.
template <int NU, int NV, int MODE, bool BT = false>
__global__ __launch_bounds__(kWave) void k_prep(const std::conditional_t<BT, BatchArgs, KArgs> args) {
    (void)args;
    using D = Dims<NU, NV, BT>;
    constexpr bool UNI = MODE == MODE_UNIFORM;
    const int E = kargs<D::BT>()->E;
    const int e = block_first_element<D::BT>() + threadIdx.x;
    const bool global = kargs<D::BT>()->cheb_global != 0;
    int n_e = 1;
    if (e < E) {
        const double* yv = kargs<D::BT>()->yv;
        double V[NV * NV];
        eload<NV * NV>(yv, E, e, V);
        n_e = stage_prep<D, UNI>(V, e, kargs<D::BT>()->serial, global);
    }
    if (global) wave_count<D::BT>(n_e, kargs<D::BT>()->slot, E);
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Volume terms (src/rhs.jl:317-333):
//   out (NO x NO) += S^T (W o (S X P^T)) P + P^T (W o (P X S^T)) S
// for the modal X (n x n) and point weights W (NQ x NQ, p + NQ q): the class
// record, one for all elements (UNI, read as scalars) or the lane's own. The
// point values of the gradient are formed one mirrored pair of point rows at a
// time and contracted back at once, so only the modal arrays stay live.
// One term of a mirrored pair of point rows (p, q' = NQ-1-p):
//   x-term (DX): out += S^T (W o (S X P^T)) P over the rows p, q'
//   y-term:      out += P^T (W o (P X S^T)) S
// The x-term reads all of P and row p of S, the y-term all of S and row p of
// P, so each holds half of the 1D tables in SGPRs.
template <class D, int n, int NO, bool DX>
__device__ __forceinline__ void vol_term(const double* tab, int p, const double (&X)[n * n], const double (&wp)[D::NQ],
                                         const double (&wq)[D::NQ], double (&out)[NO * NO]) {
    constexpr int NQ = D::NQ;
    cdbl* T = ctab(tab);
    cdbl* A = T + (DX ? D::oS : D::oP);   // contracted with X's first index, row p
    cdbl* B = T + (DX ? D::oP : D::oS);   // contracted with X's second index, all rows
    double tp[n], tq[n];
#pragma unroll
    for (int b = 0; b < n; ++b) row_pair<NQ, n, DX>(A, p, [&](int aa) { return X[aa + n * b]; }, tp[b], tq[b]);
    double zp[NO], zq[NO];
    {
        double g[NQ];
        rows_sym<NQ, n, !DX>(B, [&](int b) { return tp[b]; }, g);
#pragma unroll
        for (int q = 0; q < NQ; ++q) g[q] = wp[q] * g[q];
        cols_sym<NQ, NO, !DX>(B, g, zp);
    }
    {
        double g[NQ];
        rows_sym<NQ, n, !DX>(B, [&](int b) { return tq[b]; }, g);
#pragma unroll
        for (int q = 0; q < NQ; ++q) g[q] = wq[q] * g[q];
        cols_sym<NQ, NO, !DX>(B, g, zq);
    }
    // A[q', a] = (-1)^a A[p, a] (values), -(-1)^a A[p, a] (derivatives)
#pragma unroll
    for (int b = 0; b < NO; ++b) {
        const double zs = zp[b] + zq[b], zd = zp[b] - zq[b];
#pragma unroll
        for (int aa = 0; aa < NO; ++aa) {
            const bool same = ((aa & 1) == 0) != DX;
            out[aa + NO * b] = fma(A[p + NQ * aa], same ? zs : zd, out[aa + NO * b]);
        }
    }
}

// the middle point row (NQ odd)
template <class D, int n, int NO, bool DX>
__device__ __forceinline__ void vol_term_mid(const double* tab, const double (&X)[n * n], const double (&wm)[D::NQ],
                                             double (&out)[NO * NO]) {
    constexpr int NQ = D::NQ, m = NQ / 2;
    cdbl* T = ctab(tab);
    cdbl* A = T + (DX ? D::oS : D::oP);
    cdbl* B = T + (DX ? D::oP : D::oS);
    double t[n], z[NO], g[NQ];
#pragma unroll
    for (int b = 0; b < n; ++b) t[b] = row_mid<NQ, n, DX>(A, [&](int aa) { return X[aa + n * b]; });
    rows_sym<NQ, n, !DX>(B, [&](int b) { return t[b]; }, g);
#pragma unroll
    for (int q = 0; q < NQ; ++q) g[q] = wm[q] * g[q];
    cols_sym<NQ, NO, !DX>(B, g, z);
#pragma unroll
    for (int b = 0; b < NO; ++b)
#pragma unroll
        for (int aa = (DX ? 1 : 0); aa < NO; aa += 2) out[aa + NO * b] = fma(A[m + NQ * aa], z[b], out[aa + NO * b]);
}

// Volume terms (src/rhs.jl:317-333):
//   out (NO x NO) += S^T (W o (S X P^T)) P + P^T (W o (P X S^T)) S
// for the modal X (n x n) and the point weights W (NQ x NQ, p + NQ q) of the
// element's class (Wsrc, read as vector loads). The point values of the
// gradient are formed one mirrored pair of point rows at a time and contracted
// back at once, so only the modal arrays stay live.
template <class D, int n, int NO>
__device__ __forceinline__ void vol_pass(const double* tab, const double (&X)[n * n], const double* Wsrc,
                                         double (&out)[NO * NO]) {
    constexpr int NQ = D::NQ;
#pragma unroll 1
    for (int p = 0; p < NQ / 2; ++p) {
        double wp[NQ], wq[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) { wp[q] = Wsrc[p + NQ * q]; wq[q] = Wsrc[NQ - 1 - p + NQ * q]; }
        vol_term<D, n, NO, true>(tab, p, X, wp, wq, out);
        __builtin_amdgcn_sched_barrier(0);
        vol_term<D, n, NO, false>(tab, p, X, wp, wq, out);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (NQ & 1) {
        double wm[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) wm[q] = Wsrc[NQ / 2 + NQ * q];
        vol_term_mid<D, n, NO, true>(tab, X, wm, out);
        vol_term_mid<D, n, NO, false>(tab, X, wm, out);
    }
}

// ---------------------------------------------------------------------------
// Face slots (A, D, F) of side s of element e = j + N k (B, T, L, R) from its
// own traces (from U, V) and its neighbour's facing ones (from the neighbour's
// state, read here), the interior fluxes of src/rhs.jl:620-700; on a physical
// boundary the slots stage_boundary formed. Reduced to the side's lift vectors
// Y_A, Y_D, Y_F (src/rhs.jl:434-441), kept in the lane's LDS slots.
template <class D, bool UNI, int s>
__device__ __forceinline__ void face_side(int e, int j, int k, int N, int E, const double (&U)[D::NU * D::NU],
                                          const double (&V)[D::NV * D::NV], double* Ys, int lane) {
    constexpr int NQ = D::NQ, NV = D::NV, NU = D::NU;
    constexpr bool hz = s < 2, plus = (s == 0 || s == 2);
    constexpr int ns = s ^ 1;   // the neighbour's facing side: B <-> T, L <-> R
    CKArgs& A = *kargs<D::BT>();
    int fidx, nb, m;
    bool bnd;
    if (s == 0)      { fidx = j + (N + 1) * k;     nb = e - 1; bnd = j == 0;     m = k; }
    else if (s == 1) { fidx = j + 1 + (N + 1) * k; nb = e + 1; bnd = j == N - 1; m = k; }
    else if (s == 2) { fidx = j + N * k;           nb = e - N; bnd = k == 0;     m = j; }
    else             { fidx = j + N * (k + 1);     nb = e + N; bnd = k == N - 1; m = j; }
    if (bnd) nb = e;    // a boundary lane reads its own state here; its slots come from stage_boundary
    double odu[NQ], ov[NQ], odv[NQ], ndu[NQ], nvv[NQ], ndv[NQ];
    {
        cdbl* T = ctab(A.tab);
        side_trace<D, NU, true, s>(T, D::oE1r, U, odu);
        side_trace<D, NV, false, s>(T, D::oEr, V, ov);
        side_trace<D, NV, true, s>(T, D::oE1r, V, odv);
    }
    {
        const double* yu = A.yu;
        double X[NU * NU];
        eload<NU * NU>(yu, E, nb, X);
        side_trace<D, NU, true, ns>(ctab(A.tab), D::oE1r, X, ndu);
    }
    {
        const double* yv = A.yv;
        double X[NV * NV];
        eload<NV * NV>(yv, E, nb, X);
        cdbl* T = ctab(A.tab);
        side_trace<D, NV, false, ns>(T, D::oEr, X, nvv);
        side_trace<D, NV, true, ns>(T, D::oE1r, X, ndv);
    }
    double wA[NQ], wD[NQ], wF[NQ];
    {
        CKArgs& I = *kargs<D::BT>();
        cdbl* Tw = ctab(I.tab) + D::oW;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const long fi = (long)i * I.fstride + fidx;
            double Av, Dd, Ff;
            if (plus) interior_point<UNI>(I, true, hz, fi, nvv[i], ov[i], ndu[i], odu[i], ndv[i], odv[i], Av, Dd, Ff);
            else interior_point<UNI>(I, false, hz, fi, ov[i], nvv[i], odu[i], ndu[i], odv[i], ndv[i], Av, Dd, Ff);
            wA[i] = Tw[i] * Av;
            wD[i] = Tw[i] * Dd;
            wF[i] = Tw[i] * Ff;
        }
    }
    if (bnd) {
        CKArgs& B = *kargs<D::BT>();
        cdbl* Tw = ctab(B.tab) + D::oW;
        const double* bs = B.bsl + (long)s * NQ * 3 * N + m;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            wA[i] = Tw[i] * bs[(long)(i * 3 + 0) * N];
            wD[i] = Tw[i] * bs[(long)(i * 3 + 1) * N];
            wF[i] = Tw[i] * bs[(long)(i * 3 + 2) * N];
        }
    }
    cdbl* T = ctab(A.tab);
    double YA[NV], YD[NV], YF[NU];
    cols_sym<NQ, NV, false>(T + D::oP, wA, YA);
    cols_sym<NQ, NV, false>(T + D::oP, wD, YD);
    cols_sym<NQ, NU, false>(T + D::oP, wF, YF);
    constexpr int y0 = s * (2 * NV + NU);
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        Ys[(y0 + b) * kWave + lane] = YA[b];
        Ys[(y0 + NV + b) * kWave + lane] = YD[b];
    }
#pragma unroll
    for (int b = 0; b < NU; ++b) Ys[(y0 + 2 * NV + b) * kWave + lane] = YF[b];
}

// The lifts of side s (src/rhs.jl:449-473): Rv += (2/h) E (x) Y_A + (4/h^2) E1 (x) Y_D
// and b -= E1 (x) Y_F on a horizontal side, with the factors transposed on a
// vertical side; Y from the lane's LDS slots.
template <class D, int s>
__device__ __forceinline__ void apply_lifts(const double* Ys, int lane, double (&Rv)[D::NV * D::NV],
                                            double (&bv)[D::NU * D::NU]) {
    constexpr int NV = D::NV, NU = D::NU;
    constexpr bool hz = s < 2, plus = (s == 0 || s == 2);
    constexpr int y0 = s * (2 * NV + NU);
    double YA[NV], YD[NV], YF[NU];
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        YA[b] = Ys[(y0 + b) * kWave + lane];
        YD[b] = Ys[(y0 + NV + b) * kWave + lane];
    }
#pragma unroll
    for (int b = 0; b < NU; ++b) YF[b] = Ys[(y0 + 2 * NV + b) * kWave + lane];
    cdbl* T = ctab(kargs<D::BT>()->tab);
    constexpr int oE = plus ? D::oTEl : D::oTEr;      // (2/h) E
    constexpr int oE1 = plus ? D::oFE1l : D::oFE1r;   // (4/h^2) E1
    constexpr int oF1 = plus ? D::oE1l : D::oE1r;     // E1
#pragma unroll
    for (int b = 0; b < NV; ++b)
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) {
            double r = Rv[aa + NV * b];
            if (hz) {
                r = fma(T[oE + aa], YA[b], r);
                r = fma(T[oE1 + aa], YD[b], r);
            } else {
                r = fma(YA[aa], T[oE + b], r);
                r = fma(YD[aa], T[oE1 + b], r);
            }
            Rv[aa + NV * b] = r;
        }
#pragma unroll
    for (int b = 0; b < NU; ++b)
#pragma unroll
        for (int aa = 0; aa < NU; ++aa)
            bv[aa + NU * b] = hz ? fma(-T[oF1 + aa], YF[b], bv[aa + NU * b]) : fma(-YF[aa], T[oF1 + b], bv[aa + NU * b]);
}

// ---------------------------------------------------------------------------
// One mirrored pair of point rows (p, NH-1-p) of
//   K d = P_h^T (G o (P_h d P_h^T)) P_h
// accumulated into r; Gp, Gq are the weight rows p and NH-1-p.
template <class D>
__device__ __forceinline__ void kd_pair(cdbl* Ph, int p, const double (&d)[D::NV * D::NV],
                                        const double (&Gp)[D::NH], const double (&Gq)[D::NH],
                                        double (&r)[D::NV * D::NV]) {
    constexpr int NV = D::NV, NH = D::NH;
    double tp[NV], tq[NV];
#pragma unroll
    for (int b = 0; b < NV; ++b)
        row_pair<NH, NV, false>(Ph, p, [&](int aa) { return d[aa + NV * b]; }, tp[b], tq[b]);
    double zp[NV], zq[NV];
    {
        double y[NH];
        rows_sym<NH, NV, false>(Ph, [&](int b) { return tp[b]; }, y);
#pragma unroll
        for (int q = 0; q < NH; ++q) y[q] *= Gp[q];
        cols_sym<NH, NV, false>(Ph, y, zp);
    }
    {
        double y[NH];
        rows_sym<NH, NV, false>(Ph, [&](int b) { return tq[b]; }, y);
#pragma unroll
        for (int q = 0; q < NH; ++q) y[q] *= Gq[q];
        cols_sym<NH, NV, false>(Ph, y, zq);
    }
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        const double zs = zp[b] + zq[b], zd = zp[b] - zq[b];
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) r[aa + NV * b] = fma(Ph[p + NH * aa], (aa & 1) ? zd : zs, r[aa + NV * b]);
    }
}

// the middle point row (NH odd)
template <class D>
__device__ __forceinline__ void kd_mid(cdbl* Ph, const double (&d)[D::NV * D::NV], const double (&Gm)[D::NH],
                                       double (&r)[D::NV * D::NV]) {
    constexpr int NV = D::NV, NH = D::NH, m = NH / 2;
    double tm[NV];
#pragma unroll
    for (int b = 0; b < NV; ++b) tm[b] = row_mid<NH, NV, false>(Ph, [&](int aa) { return d[aa + NV * b]; });
    double y[NH], z[NV];
    rows_sym<NH, NV, false>(Ph, [&](int b) { return tm[b]; }, y);
#pragma unroll
    for (int q = 0; q < NH; ++q) y[q] *= Gm[q];
    cols_sym<NH, NV, false>(Ph, y, z);
#pragma unroll
    for (int b = 0; b < NV; ++b)
#pragma unroll
        for (int aa = 0; aa < NV; aa += 2) r[aa + NV * b] = fma(Ph[m + NH * aa], z[b], r[aa + NV * b]);
}

// ---------------------------------------------------------------------------
// RK4 stage update of one field (src/solver.jl:300-321) from the stage's
// derivative k of the element: stage 0 writes k itself (a plain right-hand
// side), stage 1 starts the accumulator, stages 2-3 add 2 k, stage 4 forms the
// new y. The inputs y and acc are loaded ahead of time (rk_load), so their
// transfer overlaps the computation of k.
template <int n>
__device__ __forceinline__ void rk_load(int stage, const double* y0, const double* acc, int E, int e,
                                        double (&y)[n], double (&ac)[n]) {
    if (stage >= 1) eload<n>(y0, E, e, y);
    if (stage >= 2) eload<n>(acc, E, e, ac);
}

template <int n>
__device__ __forceinline__ bool rk_store(int stage, double c, const double (&kv)[n], const double (&y)[n],
                                         const double (&ac)[n], double* acc, double* out, int E, int e) {
    bool fin = true;
    double o[n];
    if (stage == 0) {
        estore<n>(out, E, e, kv);
    } else if (stage == 1) {
#pragma unroll
        for (int m = 0; m < n; ++m) o[m] = y[m] + c * kv[m];
        estore<n>(out, E, e, o);
        estore<n>(acc, E, e, kv);
    } else if (stage < 4) {
#pragma unroll
        for (int m = 0; m < n; ++m) o[m] = y[m] + c * kv[m];
        estore<n>(out, E, e, o);
#pragma unroll
        for (int m = 0; m < n; ++m) o[m] = ac[m] + 2.0 * kv[m];
        estore<n>(acc, E, e, o);
    } else {
#pragma unroll
        for (int m = 0; m < n; ++m) {
            o[m] = y[m] + c * (ac[m] + kv[m]);
            fin = fin && isfinite(o[m]);
        }
        estore<n>(out, E, e, o);
    }
    return fin;
}

// One RK stage of one element per lane (one launch per stage): face slots and
// lifts, volume terms, source, u-equation, mass solve, the RK4 stage update,
// and the next stage's Chebyshev count of the new state (stage_prep). One wave per
// block: the lane's LDS slots hold first the lift vectors of its four sides,
// then the Chebyshev weights G = 2 beta v_q W of the first mirrored row pairs
// ([entry][lane], conflict-free); the rest of G stays in registers.
template <int NU, int NV, int MODE, bool BT = false>
__global__ __launch_bounds__(kWave) void k_stage(const std::conditional_t<BT, BatchArgs, KArgs> args) {
    (void)args;
    using D = Dims<NU, NV, BT>;
    constexpr int NQ = D::NQ, NH = D::NH, NP = D::NP, PL = D::PL, PR = D::PR;
    constexpr bool UNI = MODE == MODE_UNIFORM;
    constexpr bool FD = MODE == MODE_UNIFORM || MODE == MODE_SEPARABLE;
    constexpr bool CG = MODE == MODE_GENERAL_CG;
    constexpr int nU = NU * NU, nV = NV * NV;
    __shared__ double Gs[D::NLDS * kWave];
    if (skip_stage(*kargs<D::BT>())) return;
    const int E = kargs<D::BT>()->E, N = kargs<D::BT>()->N;
    const bool global = kargs<D::BT>()->cheb_global != 0;
    const int lane = threadIdx.x;
    const int e = block_first_element<D::BT>() + lane;
    const int nw = (E + kWave - 1) / kWave;
    // the global count, the maximum over the waves' counts of this stage: the
    // slots are read (eight at a time) before the state loads are issued and
    // reduced after, so the two latencies overlap; lanes past E load element E - 1
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
        const int ec = e < E ? e : E - 1;
        const double* yu = kargs<D::BT>()->yu;
        const double* yv = kargs<D::BT>()->yv;
        eload<nU>(yu, E, ec, U);
        eload<nV>(yv, E, ec, V);
        if (global) {
            for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
            nglob = m;
        }
    }
    int ncheb = 0, n_next = 1;
    if (e < E) {
        const int j = e % N, k = e / N;
        // the slots of the element's physical-boundary sides, from the state just loaded
        if (j == 0 || j == N - 1 || k == 0 || k == N - 1) {
            CKArgs& A = *kargs<D::BT>();
            stage_boundary<D, UNI>(U, V, e, A.serial, A.ct, A.st);
        }
        const long cc = UNI ? 0 : (long)kargs<D::BT>()->cls_ca[e];
        const double* Wc = kargs<D::BT>()->Cw + cc * NQ * NQ;
        const double* Wa = kargs<D::BT>()->Aw + cc * NQ * NQ;

        // ---- face slots of the four sides and their lift vectors (src/rhs.jl:421-473,
        // 589-776), into the lane's LDS slots
        face_side<D, UNI, 0>(e, j, k, N, E, U, V, Gs, lane);
        phase_barrier();
        face_side<D, UNI, 1>(e, j, k, N, E, U, V, Gs, lane);
        phase_barrier();
        face_side<D, UNI, 2>(e, j, k, N, E, U, V, Gs, lane);
        phase_barrier();
        face_side<D, UNI, 3>(e, j, k, N, E, U, V, Gs, lane);
        phase_barrier();

        // ---- volume terms: Rv (v-test) and VB (u-test), src/rhs.jl:317-333
        double Rv[nV], bv[nU];
#pragma unroll
        for (int m = 0; m < nV; ++m) Rv[m] = 0.0;
#pragma unroll
        for (int m = 0; m < nU; ++m) bv[m] = 0.0;
        const double v0 = V[0];
        vol_pass<D, NV, NU>(kargs<D::BT>()->tab, V, Wc, bv);    // (c^2 grad phi_u, grad v_h)
        vol_pass<D, NV, NV>(kargs<D::BT>()->tab, V, Wa, Rv);    // (alpha grad phi_v, grad v_h)
        phase_barrier();
        vol_pass<D, NU, NV>(kargs<D::BT>()->tab, U, Wc, Rv);    // (c^2 grad phi_v, grad u_h)
        phase_barrier();
        // scale and source, in the order of the reference's sum (src/rhs.jl:340-350)
        {
            const double s = -kargs<D::BT>()->four_hh;
#pragma unroll
            for (int m = 0; m < nV; ++m) Rv[m] = s * Rv[m];
            const int np = kargs<D::BT>()->nproj;
            for (int t = 0; t < np; ++t) {
                const double pc = kargs<D::BT>()->pc[t];
                const double* pr = kargs<D::BT>()->proj[t];
                double pv[nV];
                eload<nV>(pr, E, e, pv);
#pragma unroll
                for (int m = 0; m < nV; ++m) Rv[m] = Rv[m] + pc * pv[m];
            }
        }

        // the lifts, side by side B, T, L, R
        apply_lifts<D, 0>(Gs, lane, Rv, bv);
        apply_lifts<D, 1>(Gs, lane, Rv, bv);
        apply_lifts<D, 2>(Gs, lane, Rv, bv);
        apply_lifts<D, 3>(Gs, lane, Rv, bv);
        phase_barrier();
        bv[0] += v0;   // (src/rhs.jl:366-370)

        // ---- u-equation u_t = A_e^-1 b (src/rhs.jl:361-382) and the u part of
        // the RK4 stage update (src/solver.jl:300-321)
        double ku[nU];
        if (FD) {
            // A_e = c^2 (I (x) K + K (x) I) + e0 e0^T with K = Q diag(lambda) Q^T:
            // A_e^-1 b = (Q (x) Q) (Dc o ((Q (x) Q)^T b)); Q pairs modes of one parity.
            cdbl* Q = ctab(kargs<D::BT>()->tab) + D::oQ;
            double Y1[nU];
#pragma unroll
            for (int b = 0; b < NU; ++b)
#pragma unroll
                for (int i = 0; i < NU; ++i) {
                    double s = 0.0;
#pragma unroll
                    for (int aa = (i & 1); aa < NU; aa += 2) s = fma(Q[aa + NU * i], bv[aa + NU * b], s);
                    Y1[i + NU * b] = s;
                }
            const double* Dv = kargs<D::BT>()->Ainv + cc * nU;
            cdbl* Du = ctab(kargs<D::BT>()->Ainv);
#pragma unroll
            for (int i = 0; i < NU; ++i) {
                const double* Dr = UNI ? Dv : vfresh(Dv);
#pragma unroll
                for (int jj = 0; jj < NU; ++jj) {
                    double s = 0.0;
#pragma unroll
                    for (int b = (jj & 1); b < NU; b += 2) s = fma(Y1[i + NU * b], Q[b + NU * jj], s);
                    bv[i + NU * jj] = s * (UNI ? Du[i + NU * jj] : Dr[i + NU * jj]);
                }
            }
#pragma unroll
            for (int jj = 0; jj < NU; ++jj)
#pragma unroll
                for (int aa = 0; aa < NU; ++aa) {
                    double s = 0.0;
#pragma unroll
                    for (int i = (aa & 1); i < NU; i += 2) s = fma(Q[aa + NU * i], bv[i + NU * jj], s);
                    Y1[aa + NU * jj] = s;
                }
#pragma unroll
            for (int b = 0; b < NU; ++b)
#pragma unroll
                for (int aa = 0; aa < NU; ++aa) {
                    double s = 0.0;
#pragma unroll
                    for (int jj = (b & 1); jj < NU; jj += 2) s = fma(Y1[aa + NU * jj], Q[b + NU * jj], s);
                    ku[aa + NU * b] = s;
                }
        } else {
            // the stored inverse of the class, symmetric: row mm is the contiguous
            // column mm; the results pass through the lane's LDS slots (the weights
            // of the mass solve are written there only afterwards)
            const double* Ai = kargs<D::BT>()->Ainv + cc * (nU * nU);
#pragma unroll 1
            for (int mm = 0; mm < nU; ++mm) {
                const double* Ar = Ai + (long)nU * mm;
                double kv = 0.0;
#pragma unroll
                for (int jj = 0; jj < nU; ++jj) kv = fma(Ar[jj], bv[jj], kv);
                Gs[(mm % D::NLDS) * kWave + lane] = kv;
            }
#pragma unroll
            for (int mm = 0; mm < nU; ++mm) ku[mm] = Gs[(mm % D::NLDS) * kWave + lane];
        }
        {
            CKArgs& A = *kargs<D::BT>();
            if (A.recu) {
                double* ru = A.recu;
                estore<nU>(ru, E, e, ku);
            }
            double yu0[nU], acu[nU];
            rk_load<nU>(A.stage, A.y0u, A.accu, E, e, yu0, acu);
            if (!rk_store<nU>(A.stage, A.cstage, ku, yu0, acu, A.accu, A.outu, E, e)) {
                atomicMin(&kargs<D::BT>()->err->nonfinite, kargs<D::BT>()->step);
                atomicMin(&kargs<D::BT>()->err->nf_serial, kargs<D::BT>()->serial);
            }
        }

        phase_barrier();
        // ---- the mass system M(v_h) X = Rv by Chebyshev iteration
        // (src/rhs.jl:886-939)
        double Gt[PR > 0 ? 2 * PR * NH : 1];
        double Gm[NH];
        double th = 1.0, de = 1.0, s1 = 1.0;
        int n_e;
        {
            const double* yv = kargs<D::BT>()->yv;
            double V[nV];
            eload<nV>(yv, E, e, V);
            const double* betav = kargs<D::BT>()->beta + (UNI ? 0 : (long)kargs<D::BT>()->cls_b[e] * NH * NH);
            double gmin = 1.0 / 0.0, gmax = -1.0 / 0.0;
            bool nan = false;
#pragma unroll
            for (int p = 0; p < (NH + 1) / 2; ++p) {
                cdbl* T = ctab(kargs<D::BT>()->tab);
                cdbl* bu = ctab(kargs<D::BT>()->beta);
                const bool mid = (NH & 1) && p == NH / 2;
                double yp[NH], yq[NH];
                vq_rows<D>(T + D::oPh, p, mid, V, yp, yq);
#pragma unroll
                for (int q = 0; q < NH; ++q) {
                    const double gp = yp[q] * (2.0 * (UNI ? bu[p + NH * q] : betav[p + NH * q]));
                    gmin = fmin(gmin, gp);
                    gmax = fmax(gmax, gp);
                    nan = nan || (gp != gp);
                    const double Gpv = gp * T[D::oWh + p + NH * q];
                    if (mid) {
                        Gm[q] = Gpv;
                    } else {
                        const int pq = NH - 1 - p;
                        const double gq = yq[q] * (2.0 * (UNI ? bu[pq + NH * q] : betav[pq + NH * q]));
                        gmin = fmin(gmin, gq);
                        gmax = fmax(gmax, gq);
                        nan = nan || (gq != gq);
                        const double Gqv = gq * T[D::oWh + pq + NH * q];
                        if (p < PL) {
                            Gs[((2 * p) * NH + q) * kWave + lane] = Gpv;
                            Gs[((2 * p + 1) * NH + q) * kWave + lane] = Gqv;
                        } else {
                            Gt[(2 * (p - PL)) * NH + q] = Gpv;
                            Gt[(2 * (p - PL) + 1) * NH + q] = Gqv;
                        }
                    }
                }
            }
            n_e = cheb_interval(gmin, gmax, nan, kargs<D::BT>()->cheb_L, th, de, s1);
        }
        phase_barrier();
        double X[nV];
#pragma unroll
        for (int m = 0; m < nV; ++m) X[m] = 0.0;
        if constexpr (CG) {
            // conjugate gradients on M X = Rv, M = I - K, element by element
            // (src/rhs.jl:824-853): X0 = Rv, r0 = K Rv, until |r| <= tol |Rv| or
            // n_v^2 + 10 steps; a lane that has converged keeps its values.
            cdbl* Ph = ctab(kargs<D::BT>()->tab) + D::oPh;
            auto apply_K = [&](const double (&d)[nV], double (&out)[nV]) {
#pragma unroll
                for (int m = 0; m < nV; ++m) out[m] = 0.0;
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    double Gp[NH], Gq[NH];
#pragma unroll
                    for (int q = 0; q < NH; ++q) {
                        if (p < PL) {
                            Gp[q] = Gs[((2 * p) * NH + q) * kWave + lane];
                            Gq[q] = Gs[((2 * p + 1) * NH + q) * kWave + lane];
                        } else {
                            Gp[q] = Gt[(2 * (p - PL)) * NH + q];
                            Gq[q] = Gt[(2 * (p - PL) + 1) * NH + q];
                        }
                    }
                    kd_pair<D>(Ph, p, d, Gp, Gq, out);
                }
                if (NH & 1) kd_mid<D>(Ph, d, Gm, out);
            };
            const double tol = kargs<D::BT>()->mass_tol;
            double r[nV], pv[nV], Ap[nV];
#pragma unroll
            for (int m = 0; m < nV; ++m) X[m] = Rv[m];
            apply_K(X, r);
            double rr = 0.0, bb = 0.0;
#pragma unroll
            for (int m = 0; m < nV; ++m) {
                pv[m] = r[m];
                rr = fma(r[m], r[m], rr);
                bb = fma(Rv[m], Rv[m], bb);
            }
            const double r0 = sqrt(bb);
            bool act = sqrt(rr) > tol * r0, bad = false;
            const int maxit = nV + 10;   // n_v^2 + 10 (src/rhs.jl:826)
            int it = 0;
            for (; it < maxit; ++it) {
                int any = act ? 1 : 0;
                for (int off = 32; off > 0; off >>= 1) any = max(any, __shfl_xor(any, off));
                if (!any) break;
                apply_K(pv, Ap);
                double pAp = 0.0;
#pragma unroll
                for (int m = 0; m < nV; ++m) {
                    Ap[m] = pv[m] - Ap[m];
                    pAp = fma(pv[m], Ap[m], pAp);
                }
                if (act && !(pAp > 0.0)) {
                    atomicMin(&kargs<D::BT>()->err->fkey, fail_key(kargs<D::BT>()->serial, 1, e + 1));
                    act = false;
                    bad = true;
                }
                if (act) {
                    const double a = rr / pAp;
                    double rrn = 0.0;
#pragma unroll
                    for (int m = 0; m < nV; ++m) {
                        X[m] = X[m] + a * pv[m];
                        r[m] = r[m] - a * Ap[m];
                        rrn = fma(r[m], r[m], rrn);
                    }
                    const double bta = rrn / rr;
#pragma unroll
                    for (int m = 0; m < nV; ++m) pv[m] = r[m] + bta * pv[m];
                    rr = rrn;
                    act = sqrt(rr) > tol * r0;
                }
            }
            if (!bad && sqrt(rr) > tol * r0) atomicAdd(&kargs<D::BT>()->err->cg_unconverged, 1);
            ncheb = it;
        } else if (n_e == 0) {   // with the global policy stage_prep has recorded it already
            if (!global) atomicMin(&kargs<D::BT>()->err->fkey, fail_key(kargs<D::BT>()->serial, 1, e + 1));
        } else {
            const int n = global ? nglob : n_e;
            ncheb = n;
            double r[nV], d[nV];
#pragma unroll
            for (int m = 0; m < nV; ++m) { r[m] = Rv[m]; d[m] = r[m] / th; }
            double rho = 1.0 / s1;
            // the table, one scalar load for the whole iteration (the mirrored
            // half of P_h, 27 values at q = 5, stays in SGPRs)
            cdbl* Ph = ctab(kargs<D::BT>()->tab) + D::oPh;
            for (int step = 1; step <= n; ++step) {
#pragma unroll
                for (int m = 0; m < nV; ++m) X[m] += d[m];
                if (step == n) break;
                // r <- r - (d - K d)
#pragma unroll
                for (int m = 0; m < nV; ++m) r[m] = r[m] - d[m];
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    double Gp[NH], Gq[NH];
#pragma unroll
                    for (int q = 0; q < NH; ++q) {
                        if (p < PL) {
                            Gp[q] = Gs[((2 * p) * NH + q) * kWave + lane];
                            Gq[q] = Gs[((2 * p + 1) * NH + q) * kWave + lane];
                        } else {
                            Gp[q] = Gt[(2 * (p - PL)) * NH + q];
                            Gq[q] = Gt[(2 * (p - PL) + 1) * NH + q];
                        }
                    }
                    kd_pair<D>(Ph, p, d, Gp, Gq, r);
                }
                if (NH & 1) kd_mid<D>(Ph, d, Gm, r);
                const double rhon = 1.0 / (2.0 * s1 - rho);
                const double c1 = rhon * rho, c2 = 2.0 * rhon / de;
#pragma unroll
                for (int m = 0; m < nV; ++m) d[m] = c1 * d[m] + c2 * r[m];
                rho = rhon;
            }
        }
        {
            CKArgs& A = *kargs<D::BT>();
            double yv0[nV], acv[nV];
            const int stage = A.stage;
            const double c = A.cstage;
            if (A.recv) {
                double* rv = A.recv;
                estore<nV>(rv, E, e, X);
            }
            rk_load<nV>(stage, A.y0v, A.accv, E, e, yv0, acv);
            if (!rk_store<nV>(stage, c, X, yv0, acv, A.accv, A.outv, E, e)) {
                atomicMin(&kargs<D::BT>()->err->nonfinite, kargs<D::BT>()->step);
                atomicMin(&kargs<D::BT>()->err->nf_serial, kargs<D::BT>()->serial);
            }
            if (A.has_next) {
                // the new state v, as rk_store wrote it; u is read back from outu
                double Vn[nV];
#pragma unroll
                for (int m = 0; m < nV; ++m) Vn[m] = stage < 4 ? yv0[m] + c * X[m] : yv0[m] + c * (acv[m] + X[m]);
                CKArgs& B = *kargs<D::BT>();
                n_next = stage_prep<D, UNI>(Vn, e, B.serial + 1, global);
            }
        }
    }
    if (global && kargs<D::BT>()->has_next) wave_count<D::BT>(n_next, kargs<D::BT>()->slot ^ 1, E);
    for (int off = 32; off > 0; off >>= 1) ncheb = max(ncheb, __shfl_xor(ncheb, off));
    if (lane == 0 && ncheb > 0) {   // this wave's slots only: no atomics
        unsigned* ws = kargs<D::BT>()->wstat;
        const int w = block_first_element<D::BT>() / kWave;
        ws[w] += (unsigned)ncheb;
        ws[nw + w] = max(ws[nw + w], (unsigned)ncheb);
    }
}

}  // namespace wedg

namespace wedg {

// ---------------------------------------------------------------------------
// Source-mode diagnostics (src/solver.jl:856-958).

struct RecvArgs {
    const double* v;        // the state v (N, N, n_v, n_v)
    int E, nrecv;
    const int* elem;        // element of each receiver
    const double* rows;     // per receiver: P_v row at eta, then at xi ((q_v+1) each)
    double* out;            // [step][receiver]
    long step;
};

// v_h at the receivers: eta^T V xi with the basis rows at the receiver's local
// coordinates (src/solver.jl:899-905). One lane per receiver.
template <int NV>
__global__ __launch_bounds__(kWave) void k_recv(const RecvArgs a) {
    const int r = blockIdx.x * kWave + threadIdx.x;
    if (r >= a.nrecv) return;
    const int e = a.elem[r];
    const double* eta = a.rows + (long)r * 2 * NV;
    const double* xi = eta + NV;
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < NV; ++b) {
        double t = 0.0;
#pragma unroll
        for (int aa = 0; aa < NV; ++aa) t += eta[aa] * a.v[(long)(aa + NV * b) * a.E + e];
        s += t * xi[b];
    }
    a.out[a.step * a.nrecv + r] = s;
}

struct DiagArgs {
    const double* u; const double* v;
    int N, E, j0, pad;
    double h;
    const double* tab;      // the problem's table buffer
    const double* W2d;      // (nq, nq)
    const int* cls;         // class of (c^2, beta) at the standard points per element
    const double* c2;       // per class (nq, nq)
    const double* bq;
    const double* eta0;     // P_v row at the slice's local coordinate
    double* E_out; double* vmax_out; double* nd_out;   // per element, or all null
    double* slice;          // [k][i] along element row j0, or null
    double* grid;           // v_h at every point, (N nq) x (N nq) column-major with
                            // row j nq + p (y) and column k nq + q (x), or null
};

// The element values of the energy diagnostics (src/solver.jl:870-889): the
// Westervelt energy (h^2/4) sum W (1/2 c^2 |grad u_h|^2 + (1/2 - 2/3 beta v_h) v_h^2),
// max |v_h| and min (1 - 2 beta v_h) at the standard points, and on element row
// j0 the slice P_v (eta0^T V)^T (src/solver.jl:893-896), and the snapshot grid
// P_v V P_v^T of every element (src/solver.jl:919-923).
template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_diag(const DiagArgs a) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ;
    const int e = blockIdx.x * kWave + threadIdx.x;
    if (e >= a.E) return;
    const double* P = a.tab + D::oP;   // P_u (NQ x NU); P_v = its first NV columns
    const double* S = a.tab + D::oS;
    double U[NU * NU], V[NV * NV];
#pragma unroll
    for (int m = 0; m < NU * NU; ++m) U[m] = a.u[(long)m * a.E + e];
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) V[m] = a.v[(long)m * a.E + e];
    const long cc = a.cls[e];
    const double* c2 = a.c2 + cc * NQ * NQ;
    const double* bq = a.bq + cc * NQ * NQ;
    const double two_h = 2.0 / a.h;
    const int N = a.N, j = e % N, k = e / N;
    const long ld = (long)N * NQ;
    double Esum = 0.0, vmax = 0.0, nd = 1.0 / 0.0;
#pragma unroll 1
    for (int p = 0; p < NQ; ++p) {
        // row p of P U S^T, S U P^T (the gradient) and P V P^T
        double tpu[NU], tsu[NU], tpv[NV];
#pragma unroll
        for (int b = 0; b < NU; ++b) {
            double s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int aa = 0; aa < NU; ++aa) {
                s1 += P[p + NQ * aa] * U[aa + NU * b];
                s2 += S[p + NQ * aa] * U[aa + NU * b];
            }
            tpu[b] = s1;
            tsu[b] = s2;
        }
#pragma unroll
        for (int b = 0; b < NV; ++b) {
            double s = 0.0;
#pragma unroll
            for (int aa = 0; aa < NV; ++aa) s += P[p + NQ * aa] * V[aa + NV * b];
            tpv[b] = s;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double ux = 0.0, uy = 0.0, vh = 0.0;
#pragma unroll
            for (int b = 0; b < NU; ++b) {
                ux += tpu[b] * S[q + NQ * b];
                uy += tsu[b] * P[q + NQ * b];
            }
#pragma unroll
            for (int b = 0; b < NV; ++b) vh += tpv[b] * P[q + NQ * b];
            ux = two_h * ux;
            uy = two_h * uy;
            const double cq = c2[p + NQ * q], bb = bq[p + NQ * q];
            const double z = a.W2d[p + NQ * q] * (0.5 * cq * (ux * ux + uy * uy) + (0.5 - (2.0 / 3.0) * bb * vh) * (vh * vh));
            Esum += z;
            vmax = fmax(vmax, fabs(vh));
            nd = fmin(nd, 1.0 - 2.0 * bb * vh);
            if (a.grid) a.grid[(long)(j * NQ + p) + ld * (k * NQ + q)] = vh;
        }
    }
    if (a.E_out) {
        a.E_out[e] = (a.h * a.h / 4.0) * Esum;
        a.vmax_out[e] = vmax;
        a.nd_out[e] = nd;
    }
    if (a.slice && j == a.j0) {
        double t[NV];
#pragma unroll
        for (int b = 0; b < NV; ++b) {
            double s = 0.0;
#pragma unroll
            for (int aa = 0; aa < NV; ++aa) s += a.eta0[aa] * V[aa + NV * b];
            t[b] = s;
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NV; ++b) s += P[i + NQ * b] * t[b];
            a.slice[(long)k * NQ + i] = s;
        }
    }
}

}  // namespace wedg

namespace wedg {

// ---------------------------------------------------------------------------
// Error norms (src/solver.jl:774-836): per element the quadrature sums of
//   MMS:  W (u* - u_h)^2, W (v* - v_h)^2, W |grad u* - grad_h u_h|^2,
//         W (1/2 c^2 |grad u* - grad_h u_h|^2 + 1/2 (v* - v_h)^2)
//         against u* = (cos(wT) - 1) Phi, v* = -w sin(wT) Phi with
//         Phi = cos(2 pi kx x + px) sin(2 pi ky y + py), its x and y derivatives;
//   DIFF: the same sums of the difference of two states (e_u, e_v) with u* = v* = 0
//         (erru, errv, errE; the gradient sum is not used).
struct NormArgs {
    const double* u; const double* v;      // the state (MMS) or state A (DIFF)
    const double* ub; const double* vb;    // state B (DIFF), else null
    int N, E;
    double h, cu, cv, tpx, tpy;            // cos(wT) - 1, -w sin(wT), 2 pi kx, 2 pi ky
    const double* tab;
    const double* W2d;
    const int* cls; const double* c2;      // classes of c^2 at the points (wedg_problem_set_diag)
    const double* cx; const double* sx;    // cos, sin(2 pi kx x + px) at the x points [k][q]
    const double* cy; const double* sy;    // cos, sin(2 pi ky y + py) at the y points [j][p]
    double* out;                            // [4][E]
};

template <int NU, int NV>
__global__ __launch_bounds__(kWave) void k_norms(const NormArgs a) {
    using D = Dims<NU, NV>;
    constexpr int NQ = D::NQ;
    const int e = blockIdx.x * kWave + threadIdx.x;
    if (e >= a.E) return;
    const bool diff = a.ub != nullptr;
    const double* P = a.tab + D::oP;
    const double* S = a.tab + D::oS;
    double U[NU * NU], V[NV * NV];
#pragma unroll
    for (int m = 0; m < NU * NU; ++m) U[m] = a.u[(long)m * a.E + e] - (diff ? a.ub[(long)m * a.E + e] : 0.0);
#pragma unroll
    for (int m = 0; m < NV * NV; ++m) V[m] = a.v[(long)m * a.E + e] - (diff ? a.vb[(long)m * a.E + e] : 0.0);
    const int N = a.N, j = e % N, k = e / N;
    const double* c2 = a.c2 + (long)a.cls[e] * NQ * NQ;
    const double two_h = 2.0 / a.h;
    double su = 0.0, sv = 0.0, sg = 0.0, sE = 0.0;
#pragma unroll 1
    for (int p = 0; p < NQ; ++p) {
        double tpu[NU], tsu[NU], tpv[NV];
#pragma unroll
        for (int b = 0; b < NU; ++b) {
            double s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int aa = 0; aa < NU; ++aa) {
                s1 += P[p + NQ * aa] * U[aa + NU * b];
                s2 += S[p + NQ * aa] * U[aa + NU * b];
            }
            tpu[b] = s1;
            tsu[b] = s2;
        }
#pragma unroll
        for (int b = 0; b < NV; ++b) {
            double s = 0.0;
#pragma unroll
            for (int aa = 0; aa < NV; ++aa) s += P[p + NQ * aa] * V[aa + NV * b];
            tpv[b] = s;
        }
        const double syv = diff ? 0.0 : a.sy[(long)j * NQ + p], cyv = diff ? 0.0 : a.cy[(long)j * NQ + p];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double uh = 0.0, ux = 0.0, uy = 0.0, vh = 0.0;
#pragma unroll
            for (int b = 0; b < NU; ++b) {
                uh += tpu[b] * P[q + NQ * b];
                ux += tpu[b] * S[q + NQ * b];
                uy += tsu[b] * P[q + NQ * b];
            }
#pragma unroll
            for (int b = 0; b < NV; ++b) vh += tpv[b] * P[q + NQ * b];
            ux = two_h * ux;
            uy = two_h * uy;
            double eu, ev, ex, ey;
            if (diff) {
                eu = uh; ev = vh; ex = ux; ey = uy;
            } else {
                const double cxv = a.cx[(long)k * NQ + q], sxv = a.sx[(long)k * NQ + q];
                const double sp = cxv * syv;
                const double spx = -a.tpx * sxv * syv;
                const double spy = a.tpy * cxv * cyv;
                eu = a.cu * sp - uh;
                ev = a.cv * sp - vh;
                ex = a.cu * spx - ux;
                ey = a.cu * spy - uy;
            }
            const double w = a.W2d[p + NQ * q];
            const double g2 = ex * ex + ey * ey;
            su += w * (eu * eu);
            sv += w * (ev * ev);
            sg += w * g2;
            sE += w * (0.5 * c2[p + NQ * q] * g2 + 0.5 * (ev * ev));
        }
    }
    a.out[e] = su;
    a.out[(long)a.E + e] = sv;
    a.out[2L * a.E + e] = sg;
    a.out[3L * a.E + e] = sE;
}

}  // namespace wedg
