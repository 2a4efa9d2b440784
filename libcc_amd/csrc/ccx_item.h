// Shared device code of the (tuple, block-orbit) work item: constants, ordering tables,
// MFMA traits, operand-fragment streaming over the three K phases and the in-block
// transposes.  Included by the forward kernel (ccx_kernels.hip) and the adjoint
// (ccx_adjoint.hip); everything is internal linkage.
#ifndef CCX_ITEM_H
#define CCX_ITEM_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int TB = 16;                 // virtual block edge
constexpr int NWAVE = 8;               // waves per work group (2 per SIMD)
constexpr int NTHR = 64 * NWAVE;
constexpr int BPW = TB / NWAVE;        // batches (16x16 output tiles) per wave = 2
constexpr int KS = 4;                  // K per MFMA step
constexpr int CSTR_B = 17;             // padded strides of a 16^3 cube in LDS
constexpr int CSTR_A = TB * CSTR_B;    // 272
constexpr int CUBE = TB * CSTR_A;      // 4352 doubles
constexpr int VSTR = 17;
constexpr int VBLK = TB * VSTR;        // 272 doubles per staged VOE sub-block
constexpr int LDS_DOUBLES = (3 * CUBE > 36 * VBLK) ? 3 * CUBE : 36 * VBLK;

// orderings of the occupied tuple positions (0,1,2); even: 0,3,4
__device__ constexpr int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

__host__ __device__ constexpr int pair_index(int y, int z) { return y * 2 + (z > y ? z - 1 : z); }

__host__ __device__ constexpr int perm_index(int a, int b, int c)
{
    return a == 0 ? (b == 1 ? 0 : 1) : a == 1 ? (b == 0 ? 2 : 3) : (b == 0 ? 4 : 5);
}

// Term tables per phase (0: gamma=b, 1: gamma=c, 2: gamma=a) and ordering p:
// operand slot (orientation*3 + tuple position of x) and BX slot (ordered position pair).
struct Term { int s1, b1, s2, b2; };

__host__ __device__ constexpr Term term(int phase, int p)
{
    const int p1 = PERM[p][0], p2 = PERM[p][1], p3 = PERM[p][2];
    if (phase == 0)       // AX[p1](a,c) BX[p3,p2](b) ; AX[p3](c,a) BX[p1,p2](b)
        return {p1, pair_index(p3, p2), 3 + p3, pair_index(p1, p2)};
    else if (phase == 1)  // AX[p1](a,b) BX[p2,p3](c) ; AX[p2](b,a) BX[p1,p3](c)
        return {p1, pair_index(p2, p3), 3 + p2, pair_index(p1, p3)};
    else                  // AX[p2](b,c) BX[p3,p1](a) ; AX[p3](c,b) BX[p2,p1](a)
        return {p2, pair_index(p3, p1), 3 + p3, pair_index(p2, p1)};
}

// w_P at the virtual position pi(e) equals w_Q at e, with Q_n = P_{pi^-1(n)}.
__host__ __device__ constexpr int wsrc(int p, const int *pi)
{
    int inv[3] = {0, 0, 0};
    inv[pi[0]] = 0;
    inv[pi[1]] = 1;
    inv[pi[2]] = 2;
    return perm_index(PERM[p][inv[0]], PERM[p][inv[1]], PERM[p][inv[2]]);
}

__host__ __device__ constexpr int wsrc_comp(int p, int s, int t)
{
    int pi[3] = {PERM[s][PERM[t][0]], PERM[s][PERM[t][1]], PERM[s][PERM[t][2]]};
    return wsrc(p, pi);
}

// Item modes: which orderings are accumulated.  Equal blocks of the orbit (AAC: A = B, ACC:
// B = C, AAA) relate orderings by transposes inside the cube, equal occupied indices of the tuple
// ((i,j,j), (i,i,j)) pointwise; a degenerate tuple on an orbit with two equal blocks uses both.
enum Mode { M_FULL = 0, M_AAC = 1, M_ACC = 2, M_AAA = 3, M_IJJ = 4, M_IIJ = 5,
            M_IJJ_AAC = 6, M_IJJ_ACC = 7, M_IIJ_AAC = 8, M_IIJ_ACC = 9 };

__host__ __device__ constexpr int mode_mask(int mode)
{
    return mode == M_FULL ? 0x3f : mode == M_AAC ? 0x0b /*0,1,3*/ : mode == M_ACC ? 0x15 /*0,2,4*/
         : mode == M_AAA ? 0x01 : mode == M_IJJ ? 0x0d /*0,2,3*/ : mode == M_IIJ ? 0x13 /*0,1,4*/
         : mode == M_IJJ_AAC ? 0x09 /*0,3*/ : mode == M_IJJ_ACC ? 0x05 /*0,2*/
         : mode == M_IIJ_AAC ? 0x03 /*0,1*/ : 0x11 /*0,4*/;
}

// operand slots a phase needs for the computed orderings
__host__ __device__ constexpr int slot_mask(int phase, int mode)
{
    int m = 0;
    for (int p = 0; p < 6; ++p)
        if (mode_mask(mode) & (1 << p)) m |= (1 << term(phase, p).s1) | (1 << term(phase, p).s2);
    return m;
}

__host__ __device__ constexpr int bx_mask(int phase, int mode)
{
    int m = 0;
    for (int p = 0; p < 6; ++p)
        if (mode_mask(mode) & (1 << p)) m |= (1 << term(phase, p).b1) | (1 << term(phase, p).b2);
    return m;
}

template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// The two MFMA forms: operand type T, accumulator tile type, and the row a lane/item
// of a 16x16 output tile holds (column = lane & 15 for both).
template <class T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

}  // namespace

struct TriplesArgs {
    const void *axr;       // [no][nb][nb][8 alpha pairs][kp/4][4*16 lanes: k4, beta][2 alpha] (row), type T
    const void *axc;       // [no][nb][nb][8 beta pairs][kp/4][4*16 lanes: k4, alpha][2 beta] (column), type T
    const void *bx;        // [no][no][nb][kp/4][4][16 gamma], type T
    const double *voe;     // [no][no][nup][nup], voe[q][r][x][y] = VOE(x,y,q,r)
    const double *eh;      // [no]
    const double *ep;      // [nup], padded entries finite
    const int *tuples;     // [ntup][4]: i, j, k, kind (1 distinct, 2 = (i,j,j), 3 = (i,i,j))
    const int *blocks;     // [nblk][4]: A, B, C, multiplicity (1, 2 or 6)
    double *e_part;        // [nitems]
    double *v1_part;       // [nitems][3][16][3]
    double *wout;          // W_ijk cube (nu^3, column-major) of tuple 0, or null
    unsigned int *counter; // work queue head (zeroed before the launch)
    int no, nu, nup, nb, kp, ntup, nblk;
    int block_major;       // work order: 1 block triple major, 0 tuple major
    long long nitems;
    // split forms: fixed-point digit slices of AX (both orientations) and BX in the fragment
    // order of their MFMA (run_phase_split), K padded to kps (two K blocks of the form);
    // W = (g0 + g1) * wscale with wscale = sigma_A sigma_B
    const void *axr_s, *axc_s, *bx_s;
    const int *axr_e, *axc_e, *bx_e;   // FP8 form: per-lane block exponent words
    int kps;
    double wscale;
};

namespace {

struct ItemCtx {
    int tpos[3];        // occupied indices of the tuple positions
    int blk[3];         // A, B, C
    int ys[6], zs[6];   // BX slot -> ordered occupied pair
};

// Entry r (0..2) of a three-entry register array by selects.  A run-time index into the item
// context would place it in scratch memory and serialise every read behind a scratch load.
__device__ __forceinline__ int pick3(const int (&v)[3], int r) { return r == 0 ? v[0] : r == 1 ? v[1] : v[2]; }

// Fragment base pointers (wave-uniform) of one phase for this wave's two batches; a
// fragment of K step kb is base[kb*64 + lane].
template <class T>
struct PhasePtrs {
    const T *p[6];        // operand slot: the wave's two batches, interleaved per lane
    const T *x[6];        // BX slot
};

template <int PHASE, class T>
__device__ __forceinline__ void phase_ptrs(const TriplesArgs &g, const ItemCtx &it, int wave, PhasePtrs<T> &pp)
{
    const int A = it.blk[0], B = it.blk[1], C = it.blk[2];
    // orientation 0 / 1 panels (R0 = alpha block, R1 = beta block) and column-read flags
    int r0[2], r1[2];
    bool col[2];
    if (PHASE == 0) { r0[0] = A; r1[0] = C; col[0] = true;  r0[1] = C; r1[1] = A; col[1] = false; }
    else if (PHASE == 1) { r0[0] = A; r1[0] = B; col[0] = true;  r0[1] = B; r1[1] = A; col[1] = false; }
    else { r0[0] = B; r1[0] = C; col[0] = false; r0[1] = C; r1[1] = B; col[1] = true; }
    const int G = PHASE == 0 ? B : PHASE == 1 ? C : A;
    const long long per_panel = 256LL * g.kp;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int o = s / 3, xp = s % 3;
        const T *base = static_cast<const T *>(col[o] ? g.axc : g.axr) +
                        (((long long)it.tpos[xp] * g.nb + r0[o]) * g.nb + r1[o]) * per_panel;
        pp.p[s] = base + (long long)wave * (32LL * g.kp);
    }
#pragma unroll
    for (int q = 0; q < 6; ++q)
        pp.x[q] = static_cast<const T *>(g.bx) + ((long long)(it.ys[q] * g.no + it.zs[q]) * g.nb + G) * (16LL * g.kp);
}

template <class T>
struct Frags {
    T p[6][BPW];
    T x[6];
};

static_assert(BPW == 2, "a wave's two batches share one 2-wide operand load");

template <int PHASE, int MODE, class T>
__device__ __forceinline__ void load_frags(const PhasePtrs<T> &pp, int kb, int lane, Frags<T> &f)
{
    typedef typename std::conditional<std::is_same<T, double>::value, double2, float2>::type T2;
    constexpr int SM = slot_mask(PHASE, MODE), XM = bx_mask(PHASE, MODE);
    const int off = kb * 64 + lane;
#pragma unroll
    for (int s = 0; s < 6; ++s)
        if (SM & (1 << s)) {
            const T2 v = reinterpret_cast<const T2 *>(pp.p[s])[off];
            f.p[s][0] = v.x;
            f.p[s][1] = v.y;
        }
#pragma unroll
    for (int q = 0; q < 6; ++q)
        if (XM & (1 << q)) f.x[q] = pp.x[q][off];
}

template <int PHASE, int MODE, class T>
__device__ __forceinline__ void mfma_step(const Frags<T> &f, typename Mfma<T>::acc_t (&acc)[6][BPW])
{
    constexpr int MM = mode_mask(MODE);
    // first term of every computed ordering, then the second, so no MFMA waits on its predecessor
#pragma unroll
    for (int bb = 0; bb < BPW; ++bb)
        static_for<6>([&](auto P) __attribute__((always_inline)) {
            constexpr int p = decltype(P)::value;
            if constexpr ((MM >> p) & 1) {
                constexpr Term t = term(PHASE, p);
                if constexpr (PHASE == 2)
                    acc[p][bb] = Mfma<T>::mac(f.x[t.b1], f.p[t.s1][bb], acc[p][bb]);
                else
                    acc[p][bb] = Mfma<T>::mac(f.p[t.s1][bb], f.x[t.b1], acc[p][bb]);
            }
        });
#pragma unroll
    for (int bb = 0; bb < BPW; ++bb)
        static_for<6>([&](auto P) __attribute__((always_inline)) {
            constexpr int p = decltype(P)::value;
            if constexpr ((MM >> p) & 1) {
                constexpr Term t = term(PHASE, p);
                if constexpr (PHASE == 2)
                    acc[p][bb] = Mfma<T>::mac(f.x[t.b2], f.p[t.s2][bb], acc[p][bb]);
                else
                    acc[p][bb] = Mfma<T>::mac(f.p[t.s2][bb], f.x[t.b2], acc[p][bb]);
            }
        });
}

// Whether this wave's two batches reach a virtual index of the phase's batch block (C in
// phase 0, B in phases 1 and 2).  A wave whose batches lie past the last virtual index (the
// ragged last block) has only zero outputs in that phase and skips its sweep; its SIMD then
// runs the other wave's MFMAs alone.
template <int PHASE>
__device__ __forceinline__ bool phase_active(const TriplesArgs &g, const ItemCtx &it, int wave)
{
    const int bblk = PHASE == 0 ? it.blk[2] : it.blk[1];
    return wave * BPW < g.nu - bblk * TB;
}

// Three distinct blocks whose last one is ragged (fewer than 16 virtual indices): that block
// takes role B, the batch axis of two of the three K phases (role C is the batch axis of one),
// where the waves holding only padding skip the sweep.  The orbit's cube and its six
// orderings are the same sums in either role order, and every store of an item is placed by
// the item's own roles except the v1 partials, which the caller stores back in the block
// table's role order when this returns true.
__device__ __forceinline__ bool ragged_block_to_b(const TriplesArgs &g, ItemCtx &it, int mult)
{
    if (mult != 1 || g.nu - it.blk[2] * TB >= TB) return false;
    const int b = it.blk[1];
    it.blk[1] = it.blk[2];
    it.blk[2] = b;
    return true;
}

// One K sweep.  Register double buffering: fragments of step kb+1 load while step kb runs.
template <int PHASE, int MODE, class T>
__device__ __forceinline__ void run_phase(const TriplesArgs &g, const ItemCtx &it,
                                          typename Mfma<T>::acc_t (&acc)[6][BPW], int wave, int lane)
{
    if (!phase_active<PHASE>(g, it, wave)) return;
    PhasePtrs<T> pp;
    phase_ptrs<PHASE, T>(g, it, wave, pp);
    // Ring of NBUF register buffers: the loads of step kb+NBUF-1 are in flight while step kb
    // computes.  FP64 uses 2: one step covers the load latency at the FP64 MFMA rate.  Three
    // or four buffers for the items with fewer orderings, and the first fragments of the next
    // phase requested before the end of the current one, measured the same at no=24, nu=216.
    constexpr int NBUF = 2;
    const int nkb = g.kp / KS;
    const int last = nkb - 1;
    Frags<T> f[NBUF];
    static_for<NBUF - 1>([&](auto B) __attribute__((always_inline)) {
        constexpr int b = decltype(B)::value;
        load_frags<PHASE, MODE, T>(pp, b < last ? b : last, lane, f[b]);
    });
    // Every load is unconditional (prefetches past the end re-read step nkb-1), so the number
    // of loads in flight is the same on every path and each step waits only for its own
    // fragments (a conditional prefetch makes the wait counter drain completely).  The
    // scheduling barriers keep each prefetch group ahead of the current step's MFMAs (the
    // scheduler would otherwise sink loads next to their uses).
    int kb = 0;
    for (; kb + NBUF <= nkb; kb += NBUF) {
        static_for<NBUF>([&](auto J) __attribute__((always_inline)) {
            constexpr int jn = decltype(J)::value;
            const int nxt = kb + jn + NBUF - 1;
            load_frags<PHASE, MODE, T>(pp, nxt < last ? nxt : last, lane, f[(jn + NBUF - 1) % NBUF]);
            __builtin_amdgcn_sched_barrier(0);
            mfma_step<PHASE, MODE, T>(f[jn], acc);
            __builtin_amdgcn_sched_barrier(0);
        });
    }
    static_for<NBUF - 1>([&](auto J) __attribute__((always_inline)) {
        constexpr int jn = decltype(J)::value;
        if (kb + jn < nkb) mfma_step<PHASE, MODE, T>(f[jn], acc);
    });
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double group16_sum(double v)
{
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Element (a,b,c) of a 16^3 cube in LDS.
__device__ __forceinline__ int cidx(int a, int b, int c) { return a * CSTR_A + b * CSTR_B + c; }

// Layout L (after phase 0): wave owns batches b = wave*BPW+bb; lane/v give row a = (lane>>4)+4v,
// column c = lane&15.  Layout Lb (phase 0): batches c, rows a, columns b.
template <int MODE, class T>
__device__ __forceinline__ void transpose_lb_to_l(double *lds, typename Mfma<T>::acc_t (&acc)[6][BPW], int wave,
                                                  int lane)
{
    constexpr int MM = mode_mask(MODE);
    constexpr int NC = __builtin_popcount(MM);
    // up to 3 cubes per round
    static_for<2>([&](auto R) __attribute__((always_inline)) {
        constexpr int round = decltype(R)::value;
        if constexpr (round * 3 < NC) {
            int slot = 0;
            static_for<6>([&](auto P) __attribute__((always_inline)) {
                constexpr int p = decltype(P)::value;
                constexpr int rank = __builtin_popcount(MM & ((1 << p) - 1));
                if constexpr (((MM >> p) & 1) && rank / 3 == round) {
                    double *cube = lds + (rank % 3) * CUBE;
#pragma unroll
                    for (int bb = 0; bb < BPW; ++bb)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            cube[cidx(Mfma<T>::row(lane, v), lane & 15, wave * BPW + bb)] = acc[p][bb][v];
                }
            });
            (void)slot;
            __syncthreads();
            static_for<6>([&](auto P) __attribute__((always_inline)) {
                constexpr int p = decltype(P)::value;
                constexpr int rank = __builtin_popcount(MM & ((1 << p) - 1));
                if constexpr (((MM >> p) & 1) && rank / 3 == round) {
                    const double *cube = lds + (rank % 3) * CUBE;
#pragma unroll
                    for (int bb = 0; bb < BPW; ++bb)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            acc[p][bb][v] = cube[cidx(Mfma<T>::row(lane, v), wave * BPW + bb, lane & 15)];
                }
            });
            __syncthreads();
        }
    });
}

// Fill the orderings a mode did not accumulate (layout L).
template <int MODE, class T>
__device__ __forceinline__ void complete_orderings(double *lds, typename Mfma<T>::acc_t (&acc)[6][BPW], int wave,
                                                   int lane)
{
    if constexpr (MODE == M_FULL) {
        return;
    } else if constexpr (MODE == M_IJJ) {   // (i,j,j): w1 = w0, w4 = w2, w5 = w3
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb) { acc[1][bb] = acc[0][bb]; acc[4][bb] = acc[2][bb]; acc[5][bb] = acc[3][bb]; }
    } else if constexpr (MODE == M_IIJ) {   // (i,i,j): w2 = w0, w3 = w1, w5 = w4
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb) { acc[2][bb] = acc[0][bb]; acc[3][bb] = acc[1][bb]; acc[5][bb] = acc[4][bb]; }
    } else {
        // Equal blocks: w_P(e) = w_Q(pi.e) read from the cube of Q in LDS.
        constexpr int MM = mode_mask(MODE);
        static_for<6>([&](auto P) __attribute__((always_inline)) {
            constexpr int p = decltype(P)::value;
            constexpr int rank = __builtin_popcount(MM & ((1 << p) - 1));
            if constexpr ((MM >> p) & 1) {
                double *cube = lds + rank * CUBE;
#pragma unroll
                for (int bb = 0; bb < BPW; ++bb)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        cube[cidx(Mfma<T>::row(lane, v), wave * BPW + bb, lane & 15)] = acc[p][bb][v];
            }
        });
        __syncthreads();
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int e[3] = {Mfma<T>::row(lane, v), wave * BPW + bb, lane & 15};
                if constexpr (MODE == M_AAC) {         // w2(a,b,c)=w0(b,a,c), w4=w1(b,a,c), w5=w3(b,a,c)
                    acc[2][bb][v] = lds[0 * CUBE + cidx(e[1], e[0], e[2])];
                    acc[4][bb][v] = lds[1 * CUBE + cidx(e[1], e[0], e[2])];
                    acc[5][bb][v] = lds[2 * CUBE + cidx(e[1], e[0], e[2])];
                } else if constexpr (MODE == M_ACC) {  // w1=w0(a,c,b), w3=w2(a,c,b), w5=w4(a,c,b)
                    acc[1][bb][v] = lds[0 * CUBE + cidx(e[0], e[2], e[1])];
                    acc[3][bb][v] = lds[1 * CUBE + cidx(e[0], e[2], e[1])];
                    acc[5][bb][v] = lds[2 * CUBE + cidx(e[0], e[2], e[1])];
                } else if constexpr (MODE == M_IJJ_AAC) {   // cubes w0, w3: w2 = w0(b,a,c)
                    acc[2][bb][v] = lds[0 * CUBE + cidx(e[1], e[0], e[2])];
                } else if constexpr (MODE == M_IJJ_ACC) {   // cubes w0, w2: w3 = w2(a,c,b)
                    acc[3][bb][v] = lds[1 * CUBE + cidx(e[0], e[2], e[1])];
                } else if constexpr (MODE == M_IIJ_AAC) {   // cubes w0, w1: w4 = w1(b,a,c)
                    acc[4][bb][v] = lds[1 * CUBE + cidx(e[1], e[0], e[2])];
                } else if constexpr (MODE == M_IIJ_ACC) {   // cubes w0, w4: w1 = w0(a,c,b)
                    acc[1][bb][v] = lds[0 * CUBE + cidx(e[0], e[2], e[1])];
                } else {                               // A=B=C: w_P(e) = w_0(P^-1 . e)
                    static_for<5>([&](auto Q) __attribute__((always_inline)) {
                        constexpr int p = decltype(Q)::value + 1;
                        constexpr int i0 = PERM[p][0] == 0 ? 0 : PERM[p][1] == 0 ? 1 : 2;
                        constexpr int i1 = PERM[p][0] == 1 ? 0 : PERM[p][1] == 1 ? 1 : 2;
                        constexpr int i2 = PERM[p][0] == 2 ? 0 : PERM[p][1] == 2 ? 1 : 2;
                        acc[p][bb][v] = lds[cidx(e[i0], e[i1], e[i2])];
                    });
                }
            }
        // then the pointwise relations of the degenerate tuple
        if constexpr (MODE == M_IJJ_AAC || MODE == M_IJJ_ACC) {   // (i,j,j): w1 = w0, w4 = w2, w5 = w3
#pragma unroll
            for (int bb = 0; bb < BPW; ++bb) { acc[1][bb] = acc[0][bb]; acc[4][bb] = acc[2][bb]; acc[5][bb] = acc[3][bb]; }
        } else if constexpr (MODE == M_IIJ_AAC || MODE == M_IIJ_ACC) {   // (i,i,j): w2 = w0, w3 = w1, w5 = w4
#pragma unroll
            for (int bb = 0; bb < BPW; ++bb) { acc[2][bb] = acc[0][bb]; acc[3][bb] = acc[1][bb]; acc[5][bb] = acc[4][bb]; }
        }
        __syncthreads();
    }
}

// Whether form FK (0 FP64, 1 FP32, 2-4 split) runs a degenerate tuple on an orbit with equal
// blocks in the modes that use both symmetries (two orderings, or one for three equal blocks).
// Measured at no=24, nu=216: FP64 1143 -> 1132 ms, BF16x4 1375 -> 1359, FP8x4 4116 -> 4057; FP32
// and BF16x3 were no faster, so they keep the three-ordering modes (and their code).
__host__ __device__ constexpr bool pair_modes(int fk) { return fk != 1 && fk != 2; }

// The mode of an item (wave-uniform) in the forms with pair_modes: f is called with
// std::integral_constant<int, MODE>.  An orbit of three equal blocks accumulates one ordering
// whatever the tuple; a degenerate tuple on an orbit with two equal blocks accumulates two (the
// third of its three distinct orderings is a transpose of one of them).  The other forms select
// the modes by an explicit chain at the call site (tkind 2 / 3 first: M_IJJ / M_IIJ on any
// orbit); passing that chain through this helper's lambda changed their code generation
// (scratch 260 -> 288 B/lane for FP32) and cost 2-3 %.
template <class F>
__device__ __forceinline__ void dispatch_mode(int tkind, int mult, const ItemCtx &it, F &&f)
{
    const bool aac = it.blk[0] == it.blk[1];
    if (tkind == 2) {
        if (mult == 1) f(std::integral_constant<int, M_IJJ>{});
        else if (mult == 6) f(std::integral_constant<int, M_AAA>{});
        else if (aac) f(std::integral_constant<int, M_IJJ_AAC>{});
        else f(std::integral_constant<int, M_IJJ_ACC>{});
    } else if (tkind == 3) {
        if (mult == 1) f(std::integral_constant<int, M_IIJ>{});
        else if (mult == 6) f(std::integral_constant<int, M_AAA>{});
        else if (aac) f(std::integral_constant<int, M_IIJ_AAC>{});
        else f(std::integral_constant<int, M_IIJ_ACC>{});
    } else {
        if (mult == 1) f(std::integral_constant<int, M_FULL>{});
        else if (mult == 6) f(std::integral_constant<int, M_AAA>{});
        else if (aac) f(std::integral_constant<int, M_AAC>{});
        else f(std::integral_constant<int, M_ACC>{});
    }
}

// Work-item context of tuple t and block triple q (wave-uniform values).
__device__ __forceinline__ ItemCtx make_item_ctx(const int *tuples, const int *blocks, int t, int q, int &tkind,
                                                 int &mult)
{
    int4 tp = reinterpret_cast<const int4 *>(tuples)[t];
    int4 bp = reinterpret_cast<const int4 *>(blocks)[q];
    tp.x = __builtin_amdgcn_readfirstlane(tp.x); tp.y = __builtin_amdgcn_readfirstlane(tp.y);
    tp.z = __builtin_amdgcn_readfirstlane(tp.z); tp.w = __builtin_amdgcn_readfirstlane(tp.w);
    bp.x = __builtin_amdgcn_readfirstlane(bp.x); bp.y = __builtin_amdgcn_readfirstlane(bp.y);
    bp.z = __builtin_amdgcn_readfirstlane(bp.z); bp.w = __builtin_amdgcn_readfirstlane(bp.w);
    ItemCtx it;
    it.tpos[0] = tp.x; it.tpos[1] = tp.y; it.tpos[2] = tp.z;
    it.blk[0] = bp.x; it.blk[1] = bp.y; it.blk[2] = bp.z;
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int z = 0; z < 3; ++z)
            if (y != z) { it.ys[pair_index(y, z)] = it.tpos[y]; it.zs[pair_index(y, z)] = it.tpos[z]; }
    tkind = tp.w;
    mult = bp.w;
    return it;
}

// The W of an item: the three K phases and the completion of the orderings its mode
// does not accumulate.  On return acc[p][bb][v] = w_p at (row(lane,v), wave*BPW+bb, lane&15).
template <int MODE, class T>
__device__ __forceinline__ void item_w(const TriplesArgs &g, double *lds, const ItemCtx &it,
                                       typename Mfma<T>::acc_t (&acc)[6][BPW], int wave, int lane)
{
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb) acc[p][bb] = typename Mfma<T>::acc_t{0, 0, 0, 0};
    run_phase<0, MODE, T>(g, it, acc, wave, lane);
    transpose_lb_to_l<MODE, T>(lds, acc, wave, lane);
    run_phase<1, MODE, T>(g, it, acc, wave, lane);
    run_phase<2, MODE, T>(g, it, acc, wave, lane);
    complete_orderings<MODE, T>(lds, acc, wave, lane);
}

// ---- tables shared by the LDS-staged sweeps --------------------------------------------
// n-th set bit of a mask (compile time)
__host__ __device__ constexpr int nth_bit(int mask, int n)
{
    for (int b = 0; b < 32; ++b)
        if ((mask >> b) & 1) {
            if (n == 0) return b;
            --n;
        }
    return -1;
}

// Terms of one phase that read AX slot s: ordering p and BX operand b of each.
struct SlotTerms {
    int n;
    int p[12], b[12];
};
__host__ __device__ constexpr SlotTerms slot_terms(int phase, int mode, int s)
{
    SlotTerms r{0, {}, {}};
    const int mm = mode_mask(mode);
    for (int p = 0; p < 6; ++p)
        if ((mm >> p) & 1) {
            const Term t = term(phase, p);
            if (t.s1 == s) { r.p[r.n] = p; r.b[r.n] = t.b1; ++r.n; }
            if (t.s2 == s) { r.p[r.n] = p; r.b[r.n] = t.b2; ++r.n; }
        }
    return r;
}

// ---- split K sweeps ------------------------------------------------------------------
// Each FP64 operand element is carried as NS low-precision slices whose sum reproduces it; the
// slice products of levels i + j <= NS - 1 run on the BF16 or FP8 matrix cores with FP32
// accumulation, and W is formed in FP64.  Three slicings (SplitCfg):
//   FK 2, BF16 x 3, per-element: slice s is the bf16 rounding of what slices 0..s-1 left, so
//     every element keeps 24 bits relative to its own magnitude; one FP32 accumulator.
//   FK 3, FP8 e4m3 x 4, per-row block scaled: each 32-K group of a row (one lane's fragment) has
//     its own power-of-two exponent e; x / 2^e = sum_s d_s 2^-4(s+1) with integer digits
//     |d_0| <= 16, |d_s| <= 8 stored as e4m3, and the slice weight 2^(e - 4(s+1)) rides on the
//     E8M0 block scales of v_mfma_scale_f32_16x16x128_f8f6f4 (a per-lane word holds the four
//     biased slice exponents, selected by the MFMA's byte select); one FP32 accumulator.
//   FK 4, BF16 x 4, fixed point under one power-of-two scale per slab: slice s = d_s 2^-8(s+1)
//     (|d_0| <= 256, |d_s| <= 128); the level-0 products all lie on the grid 2^-16 sigma_A
//     sigma_B and go to their own FP32 accumulator, which adds them exactly while its value
//     stays below 2^8 sigma_A sigma_B; the higher levels (each at most 2^-7 of level 0) go to a
//     second one.  W = sigma_A sigma_B (g0 + g1).
template <int FK> struct SplitCfg {
    static_assert(FK >= 2 && FK <= 4, "split form");
    static constexpr int FMT = FK == 3 ? 1 : 0;       // 0 bf16, 1 fp8 e4m3
    static constexpr int NS = FK == 2 ? 3 : 4;        // slices
    static constexpr bool TWO_ACC = FK == 4;          // level 0 in its own accumulator
    static constexpr bool MX = FK == 3;               // per-lane block exponents
    static constexpr int KB = FMT == 0 ? 32 : 128;    // K per MFMA
    static constexpr int F16 = FMT == 0 ? 1 : 2;      // 16-byte units per lane per fragment
};

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef const __attribute__((address_space(1))) i32x4 g16;   // global memory, 16 B per lane
typedef const __attribute__((address_space(1))) int g4;      // global memory, 4 B per lane

// One lane's operand fragment: 16 B (bf16 x 8) or 32 B (e4m3 x 32), loaded as 16-byte units.
template <int FK> using split_frag_t = typename std::conditional<SplitCfg<FK>::FMT == 0, i32x4, i32x8>::type;

template <int FK>
__device__ __forceinline__ split_frag_t<FK> split_frag_load(g16 *p, long long unit)
{
    if constexpr (SplitCfg<FK>::FMT == 0) {
        return p[unit];
    } else {
        const i32x4 lo = p[unit], hi = p[unit + 64];   // the lane's two 16-byte units
        return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    }
}

template <int FK>
__device__ __forceinline__ split_frag_t<FK> split_frag_lds(const i32x4 *p, int unit)
{
    if constexpr (SplitCfg<FK>::FMT == 0) {
        return p[unit];
    } else {
        const i32x4 lo = p[unit], hi = p[unit + 64];
        return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
    }
}

template <int FK, int SI, int SJ>
__device__ __forceinline__ f4 split_mac(const split_frag_t<FK> &a, const split_frag_t<FK> &b, int ea, int eb, f4 c)
{
    if constexpr (SplitCfg<FK>::FMT == 0) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    } else {
        // operand format 0 (e4m3) for A and B; byte SI of ea and byte SJ of eb are the E8M0 scales
        return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, SI, ea, SJ, eb);
    }
}

// An empty asm that reads and writes the accumulators of the orderings MODE computes: every
// MFMA into them must be issued before it, and it stays where it is among the side effects
// (barriers) around it.
template <int MODE, int N>
__device__ __forceinline__ void pin_accumulators(f4 (&acc)[N][BPW])
{
#pragma unroll
    for (int p = 0; p < N; ++p)
        if ((mode_mask(MODE) >> p) & 1)
#pragma unroll
            for (int bb = 0; bb < BPW; ++bb) asm volatile("" : "+v"(acc[p][bb]));
}

// Blocks of KB K: the BX slices of the phase (shared by all waves) are staged in LDS, double
// buffered across blocks; the AX slices stream per operand slot through a two-deep register
// ring (two blocks per loop iteration keep the ring index a compile-time constant).  A slot
// step walks the BX slices from the last to the first, reads each from LDS just before its
// products (so one BX slice is live at a time) and issues the products with every AX slice
// whose level fits, for every term and batch; consecutive MFMAs go to different accumulators.
template <int PHASE, int MODE, int FK>
__device__ __forceinline__ void run_phase_split(const TriplesArgs &g, const ItemCtx &it, f4 (&g0)[6][BPW],
                                                f4 (&g1)[SplitCfg<FK>::TWO_ACC ? 6 : 1][BPW], i32x4 *lbx,
                                                int wave, int lane)
{
    typedef SplitCfg<FK> CF;
    constexpr int SM = slot_mask(PHASE, MODE), XM = bx_mask(PHASE, MODE);
    constexpr int NSL = __builtin_popcount(SM), NBX = __builtin_popcount(XM);
    constexpr int NS = CF::NS, F16 = CF::F16;
    constexpr int CH = NS * F16;               // 16-byte chunks per lane per fragment set
    constexpr int NCH = NBX * CH;              // BX chunks per K block
    static_assert(2 * NCH * 1024 + 2 * NBX * 256 <= (int)(sizeof(double) * LDS_DOUBLES), "BX staging fits LDS");
    const int A = it.blk[0], B = it.blk[1], C = it.blk[2];
    int r0[2], r1[2];
    bool col[2];
    if (PHASE == 0) { r0[0] = A; r1[0] = C; col[0] = true;  r0[1] = C; r1[1] = A; col[1] = false; }
    else if (PHASE == 1) { r0[0] = A; r1[0] = B; col[0] = true;  r0[1] = B; r1[1] = A; col[1] = false; }
    else { r0[0] = B; r1[0] = C; col[0] = false; r0[1] = C; r1[1] = B; col[1] = true; }
    const int G = PHASE == 0 ? B : PHASE == 1 ? C : A;
    const int nkb = g.kps / CF::KB;
    // units of 16 B: an AX panel is 16 outer rows x nkb blocks x CH chunks x 64 lanes, a BX panel
    // nkb x CH x 64; this wave's outer rows are 2 wave and 2 wave + 1.  Block exponents (MX): one
    // int per lane and K block, [panel][outer][kb][lane] and [pair][G][kb][lane].
    g16 *ap[6];
    g4 *aep[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int o = s / 3, xp = s % 3;
        const long long panel = ((long long)it.tpos[xp] * g.nb + r0[o]) * g.nb + r1[o];
        ap[s] = (g16 *)(col[o] ? g.axc_s : g.axr_s) + (panel * 16 + 2 * wave) * ((long long)nkb * CH * 64);
        if constexpr (CF::MX) aep[s] = (g4 *)(col[o] ? g.axc_e : g.axr_e) + (panel * 16 + 2 * wave) * ((long long)nkb * 64);
    }
    g16 *xq[6];
    g4 *xe[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const long long pg = (long long)(it.ys[q] * g.no + it.zs[q]) * g.nb + G;
        xq[q] = (g16 *)g.bx_s + pg * ((long long)nkb * CH * 64);
        if constexpr (CF::MX) xe[q] = (g4 *)g.bx_e + pg * ((long long)nkb * 64);
    }
    int *lbe = reinterpret_cast<int *>(lbx + 2 * NCH * 64);   // BX block exponents, [buf][rank][lane]

    // BX staging: chunk c = rank * CH + (slice * F16 + f) of K block kb goes straight from global
    // memory into LDS buffer buf (global_load_lds_dwordx4: lane l's 16 bytes land at the chunk's
    // base + 16 l), loaded by wave c % NWAVE (wave w takes chunks w, w + NWAVE, ...); the
    // exponent word of rank q by wave q.  No registers hold the staged data.  The rank of a chunk
    // is wave-uniform, so its panel pointer is a uniform select.  Completion is waited for
    // (vmcnt 0) before the barrier that publishes the buffer.
    typedef __attribute__((address_space(3))) void lds_void;
    typedef const __attribute__((address_space(1))) void g_void;
    constexpr int NBR = (NCH + NWAVE - 1) / NWAVE;
    auto bx_src = [&](int rank) __attribute__((always_inline)) {
        g16 *p = xq[nth_bit(XM, 0)];
        static_for<NBX>([&](auto R) __attribute__((always_inline)) {
            constexpr int r = decltype(R)::value;
            if (rank == r) p = xq[nth_bit(XM, r)];
        });
        return p;
    };
    auto bx_load = [&](int kb, int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < NBR; ++r) {
            const int c = wave + NWAVE * r;
            if (c < NCH)
                __builtin_amdgcn_global_load_lds((g_void *)(bx_src(c / CH) + ((long long)kb * CH + c % CH) * 64 + lane),
                                                 (lds_void *)(lbx + (buf * NCH + c) * 64), 16, 0, 0);
        }
        if constexpr (CF::MX) {
            if (wave < NBX) {
                g4 *e = xe[nth_bit(XM, 0)];
                static_for<NBX>([&](auto R) __attribute__((always_inline)) {
                    constexpr int r = decltype(R)::value;
                    if (wave == r) e = xe[nth_bit(XM, r)];
                });
                __builtin_amdgcn_global_load_lds((g_void *)(e + (long long)kb * 64 + lane),
                                                 (lds_void *)(lbe + (buf * NBX + wave) * 64), 4, 0, 0);
            }
        }
    };
    // A ring step is SPS operand slots and BS of the wave's batches: one slot of both batches.
    // Two bf16 slots per step (a longer prefetch distance) measured 4 % (BF16x3) and 12 %
    // (BF16x4) slower at no=24, nu=216: the larger ring costs registers, and the sweep is not
    // waiting on its loads.  The 32-byte fp8 fragments took one batch per step while the
    // compiler spilled the sweep; with the MFMAs pinned to their block (pin_accumulators) both
    // batches fit: 879 ms against 1271 ms.
    typedef split_frag_t<FK> frag_t;
    constexpr int BS = BPW, NH = BPW / BS;
    constexpr int SPS = 1;
    constexpr int NG = (NSL + SPS - 1) / SPS;          // slot groups per K block
    frag_t ab[2][SPS][BS][NS];
    int ae[2][SPS][BS];
    auto a_load = [&](auto Gc, int h, int kb, frag_t (&d)[SPS][BS][NS], int (&de)[SPS][BS]) __attribute__((always_inline)) {
        constexpr int gi = decltype(Gc)::value;
        static_for<SPS>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = decltype(U)::value, j = gi * SPS + u;
            if constexpr (j < NSL) {
                constexpr int s = nth_bit(SM, j);
#pragma unroll
                for (int b = 0; b < BS; ++b) {
                    const int bb = h * BS + b;
#pragma unroll
                    for (int sl = 0; sl < NS; ++sl)
                        d[u][b][sl] = split_frag_load<FK>(ap[s], (((long long)bb * nkb + kb) * CH + sl * F16) * 64 + lane);
                    if constexpr (CF::MX) de[u][b] = aep[s][((long long)bb * nkb + kb) * 64 + lane];
                }
            }
        });
    };
    auto compute_slot = [&](auto Sc, auto Hc, const frag_t (&a)[BS][NS], const int (&ea)[BS], int buf)
                            __attribute__((always_inline)) {
        constexpr int s = decltype(Sc)::value, h = decltype(Hc)::value;
        constexpr SlotTerms ST = slot_terms(PHASE, MODE, s);
        int eb[ST.n > 0 ? ST.n : 1];
        static_for<ST.n>([&](auto N) __attribute__((always_inline)) {
            constexpr int n = decltype(N)::value;
            constexpr int rank = __builtin_popcount(XM & ((1 << ST.b[n]) - 1));
            eb[n] = CF::MX ? lbe[(buf * NBX + rank) * 64 + lane] : 0;
        });
        static_for<NS>([&](auto SJ) __attribute__((always_inline)) {
            constexpr int sj = NS - 1 - decltype(SJ)::value;     // BX slice, last first
            // A memory clobber per BX slice: without it the compiler merges the LDS reads of the
            // same slice across the steps of a block and keeps them all live (for the fp8 slices
            // that alone exceeds the register file); each step re-reads its slice instead.
            asm volatile("" ::: "memory");
            frag_t bs[ST.n > 0 ? ST.n : 1];
            static_for<ST.n>([&](auto N) __attribute__((always_inline)) {
                constexpr int n = decltype(N)::value;
                constexpr int rank = __builtin_popcount(XM & ((1 << ST.b[n]) - 1));
                bs[n] = split_frag_lds<FK>(lbx, (buf * NCH + rank * CH + sj * F16) * 64 + lane);
            });
            static_for<NS - sj>([&](auto SI) __attribute__((always_inline)) {
                constexpr int si = NS - 1 - sj - decltype(SI)::value;   // AX slice, highest level first
                static_for<ST.n>([&](auto N) __attribute__((always_inline)) {
                    constexpr int n = decltype(N)::value;
                    constexpr int p = ST.p[n];
#pragma unroll
                    for (int b = 0; b < BS; ++b) {
                        const int bb = h * BS + b;
                        f4 &acc = (CF::TWO_ACC && si + sj > 0) ? g1[CF::TWO_ACC ? p : 0][bb] : g0[p][bb];
                        // PHASE 2 tiles are produced transposed (BX rows, AX columns)
                        if constexpr (PHASE == 2) acc = split_mac<FK, sj, si>(bs[n], a[b][si], eb[n], ea[b], acc);
                        else acc = split_mac<FK, si, sj>(a[b][si], bs[n], ea[b], eb[n], acc);
                    }
                });
            });
        });
    };
    auto compute = [&](auto Gc, auto Hc, const frag_t (&a)[SPS][BS][NS], const int (&ea)[SPS][BS], int buf)
                       __attribute__((always_inline)) {
        constexpr int gi = decltype(Gc)::value;
        static_for<SPS>([&](auto U) __attribute__((always_inline)) {
            constexpr int u = decltype(U)::value, j = gi * SPS + u;
            if constexpr (j < NSL) compute_slot(std::integral_constant<int, nth_bit(SM, j)>{}, Hc, a[u], ea[u], buf);
        });
    };

    // steps of one K block: (slot group g, batch group h), g major; NST = NG * NH per block
    constexpr int NST = NG * NH;
    // s_waitcnt immediate (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
    // that waits until at most NPF vector-memory loads are outstanding, nothing else
    constexpr int NPF = (NSL < SPS ? NSL : SPS) * BS * (NS * F16 + (CF::MX ? 1 : 0));
    static_assert(NPF < 64, "vmcnt range");
    constexpr int WAIT_DMA = (NPF & 15) | ((NPF >> 4) << 14) | (7 << 4) | (15 << 8);
    bx_load(0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (!phase_active<PHASE>(g, it, wave)) {
        // only padding batches: stage this wave's BX chunks and meet every barrier
        for (int kb = 0; kb < nkb; kb += 2) {
            static_for<2>([&](auto HB) __attribute__((always_inline)) {
                constexpr int hb = decltype(HB)::value;
                const int kbb = kb + hb;
                bx_load(kbb + 1 < nkb ? kbb + 1 : nkb - 1, 1 - hb);
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            });
        }
        return;
    }
    a_load(std::integral_constant<int, 0>{}, 0, 0, ab[0], ae[0]);
    for (int kb = 0; kb < nkb; kb += 2) {
        static_for<2>([&](auto HB) __attribute__((always_inline)) {
            constexpr int hb = decltype(HB)::value;
            const int kbb = kb + hb;
            // buffer 1 - hb was last read in the previous block, which ended with a barrier
            bx_load(kbb + 1 < nkb ? kbb + 1 : nkb - 1, 1 - hb);
            static_for<NST>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                constexpr int t = hb * NST + j;
                constexpr int gi = j / NH, h = j % NH;
                constexpr int gn = ((j + 1) % NST) / NH, hn = ((j + 1) % NST) % NH;
                const int kn = j + 1 < NST ? kbb : (kbb + 1 < nkb ? kbb + 1 : nkb - 1);
                a_load(std::integral_constant<int, gn>{}, hn, kn, ab[(t + 1) % 2], ae[(t + 1) % 2]);
                __builtin_amdgcn_sched_barrier(0);
                compute(std::integral_constant<int, gi>{}, std::integral_constant<int, h>{}, ab[t % 2], ae[t % 2], hb);
                __builtin_amdgcn_sched_barrier(0);
            });
            // the staged block must have landed before the barrier publishes it; the AX prefetch
            // of the next step (the NPF most recent loads; loads retire in issue order) may stay
            // in flight across the barrier
            // the block's MFMAs stay before the barrier: without this the compiler may sink
            // them past it, after the next block's loads, and keep both blocks' operands live
            pin_accumulators<MODE>(g0);
            if constexpr (CF::TWO_ACC) pin_accumulators<MODE>(g1);
            __builtin_amdgcn_s_waitcnt(WAIT_DMA);
            __syncthreads();
        });
    }
}

// W of an item in a split form: g0 (and g1 for the two-accumulator form) on return hold
// w_p = (g0 + g1) * wscale at (row(lane,v), wave*BPW+bb, lane&15) in the FP32 MFMA layout.
template <int MODE, int FK>
__device__ __forceinline__ void item_w_split(const TriplesArgs &g, double *lds, const ItemCtx &it,
                                             f4 (&g0)[6][BPW], f4 (&g1)[SplitCfg<FK>::TWO_ACC ? 6 : 1][BPW],
                                             int wave, int lane)
{
    constexpr bool TWO = SplitCfg<FK>::TWO_ACC;
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb) {
            g0[p][bb] = f4{0, 0, 0, 0};
            if constexpr (TWO) g1[p][bb] = f4{0, 0, 0, 0};
        }
    i32x4 *lbx = reinterpret_cast<i32x4 *>(lds);
    run_phase_split<0, MODE, FK>(g, it, g0, g1, lbx, wave, lane);
    transpose_lb_to_l<MODE, float>(lds, g0, wave, lane);
    if constexpr (TWO) transpose_lb_to_l<MODE, float>(lds, g1, wave, lane);
    run_phase_split<1, MODE, FK>(g, it, g0, g1, lbx, wave, lane);
    run_phase_split<2, MODE, FK>(g, it, g0, g1, lbx, wave, lane);
    complete_orderings<MODE, float>(lds, g0, wave, lane);
    if constexpr (TWO) complete_orderings<MODE, float>(lds, g1, wave, lane);
}

// ---- FP32 K sweep -------------------------------------------------------------------
typedef const __attribute__((address_space(1))) f4 gf4;   // global memory, 16 B per lane

// FP32 form, blocks of 16 K: one 16-byte load per lane carries four v_mfma_f32_16x16x4_f32
// k-steps of one batch row (lane l: inner index l & 15, float j at K = 16 kb + 4 j + (l >> 4)).
// The BX fragments of the phase are staged in LDS (one 1 KB chunk per BX operand and block,
// three buffers); the AX fragments stream per operand slot through a three-deep register ring
// that runs two slot steps ahead.  Three blocks per loop iteration (K padded to 48) keep every
// ring and buffer index a compile-time constant.  4 bytes per operand element instead of 8.
template <int PHASE, int MODE>
__device__ __forceinline__ void run_phase_f32v(const TriplesArgs &g, const ItemCtx &it, f4 (&acc)[6][BPW],
                                               f4 *lbx, int wave, int lane)
{
    constexpr int SM = slot_mask(PHASE, MODE), XM = bx_mask(PHASE, MODE);
    constexpr int NSL = __builtin_popcount(SM), NBX = __builtin_popcount(XM);
    static_assert(NBX <= NWAVE, "one BX chunk per wave");
    const int A = it.blk[0], B = it.blk[1], C = it.blk[2];
    int r0[2], r1[2];
    bool col[2];
    if (PHASE == 0) { r0[0] = A; r1[0] = C; col[0] = true;  r0[1] = C; r1[1] = A; col[1] = false; }
    else if (PHASE == 1) { r0[0] = A; r1[0] = B; col[0] = true;  r0[1] = B; r1[1] = A; col[1] = false; }
    else { r0[0] = B; r1[0] = C; col[0] = false; r0[1] = C; r1[1] = B; col[1] = true; }
    const int G = PHASE == 0 ? B : PHASE == 1 ? C : A;
    const int nkb = g.kps / 16;
    // units of 16 B: an AX panel is 16 outer rows x nkb blocks x 64 lanes, a BX panel nkb x 64
    gf4 *ap[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int o = s / 3, xp = s % 3;
        const long long panel = ((long long)it.tpos[xp] * g.nb + r0[o]) * g.nb + r1[o];
        ap[s] = (gf4 *)(col[o] ? g.axc : g.axr) + panel * (16LL * nkb * 64) + (long long)wave * (2 * nkb * 64);
    }
    gf4 *xq[6];
#pragma unroll
    for (int q = 0; q < 6; ++q)
        xq[q] = (gf4 *)g.bx + ((long long)(it.ys[q] * g.no + it.zs[q]) * g.nb + G) * ((long long)nkb * 64);

    f4 br;
    auto bx_load = [&](int kb) __attribute__((always_inline)) {
        static_for<NBX>([&](auto Cc) __attribute__((always_inline)) {
            constexpr int c = decltype(Cc)::value;
            if (wave == c) br = xq[nth_bit(XM, c)][kb * 64 + lane];
        });
    };
    auto bx_store = [&](int buf) __attribute__((always_inline)) {
        static_for<NBX>([&](auto Cc) __attribute__((always_inline)) {
            constexpr int c = decltype(Cc)::value;
            if (wave == c) lbx[(buf * NBX + c) * 64 + lane] = br;
        });
    };
    f4 ab[3][BPW];
    auto a_load = [&](int s, int kb, f4 (&d)[BPW]) __attribute__((always_inline)) {
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb) d[bb] = ap[s][(bb * nkb + kb) * 64 + lane];
    };
    auto compute = [&](auto Sc, const f4 (&a)[BPW], int buf) __attribute__((always_inline)) {
        constexpr int s = decltype(Sc)::value;
        constexpr SlotTerms ST = slot_terms(PHASE, MODE, s);
        f4 bs[ST.n > 0 ? ST.n : 1];
        static_for<ST.n>([&](auto N) __attribute__((always_inline)) {
            constexpr int n = decltype(N)::value;
            constexpr int rank = __builtin_popcount(XM & ((1 << ST.b[n]) - 1));
            bs[n] = lbx[(buf * NBX + rank) * 64 + lane];
        });
        static_for<4>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            static_for<ST.n>([&](auto N) __attribute__((always_inline)) {
                constexpr int n = decltype(N)::value;
                constexpr int p = ST.p[n];
#pragma unroll
                for (int bb = 0; bb < BPW; ++bb) {
                    if constexpr (PHASE == 2)
                        acc[p][bb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bs[n][j], a[bb][j], acc[p][bb], 0, 0, 0);
                    else
                        acc[p][bb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[bb][j], bs[n][j], acc[p][bb], 0, 0, 0);
                }
            });
        });
    };

    bx_load(0);
    bx_store(0);
    __syncthreads();
    if (!phase_active<PHASE>(g, it, wave)) {
        // only padding batches: stage this wave's BX chunks and meet every barrier
        for (int kb = 0; kb < nkb; kb += 3) {
            static_for<3>([&](auto HB) __attribute__((always_inline)) {
                constexpr int hb = decltype(HB)::value;
                const int kbb = kb + hb;
                bx_load(kbb + 1 < nkb ? kbb + 1 : nkb - 1);
                bx_store((hb + 1) % 3);
                __syncthreads();
            });
        }
        return;
    }
    // steps t = block * NSL + slot index; the ring holds steps t, t+1, t+2
    a_load(nth_bit(SM, 0), 0, ab[0]);
    a_load(nth_bit(SM, 1 % NSL), 1 / NSL < nkb ? 1 / NSL : nkb - 1, ab[1]);
    for (int kb = 0; kb < nkb; kb += 3) {
        static_for<3>([&](auto HB) __attribute__((always_inline)) {
            constexpr int hb = decltype(HB)::value;
            const int kbb = kb + hb;
            bx_load(kbb + 1 < nkb ? kbb + 1 : nkb - 1);
            static_for<NSL>([&](auto J) __attribute__((always_inline)) {
                constexpr int j = decltype(J)::value;
                constexpr int t = hb * NSL + j;
                constexpr int s = nth_bit(SM, j);
                constexpr int t2 = t + 2;                       // the step loaded now
                constexpr int s2 = nth_bit(SM, t2 % NSL);
                const int k2 = kb + t2 / NSL;
                a_load(s2, k2 < nkb ? k2 : nkb - 1, ab[t2 % 3]);
                __builtin_amdgcn_sched_barrier(0);
                compute(std::integral_constant<int, s>{}, ab[t % 3], hb);
                __builtin_amdgcn_sched_barrier(0);
            });
            bx_store((hb + 1) % 3);
            __syncthreads();
        });
    }
}

template <int MODE>
__device__ __forceinline__ void item_w_f32v(const TriplesArgs &g, double *lds, const ItemCtx &it,
                                            f4 (&acc)[6][BPW], int wave, int lane)
{
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int bb = 0; bb < BPW; ++bb) acc[p][bb] = f4{0, 0, 0, 0};
    f4 *lbx = reinterpret_cast<f4 *>(lds);
    run_phase_f32v<0, MODE>(g, it, acc, lbx, wave, lane);
    transpose_lb_to_l<MODE, float>(lds, acc, wave, lane);
    run_phase_f32v<1, MODE>(g, it, acc, lbx, wave, lane);
    run_phase_f32v<2, MODE>(g, it, acc, lbx, wave, lane);
    complete_orderings<MODE, float>(lds, acc, wave, lane);
}


}  // namespace

#endif  // CCX_ITEM_H
