// Closed-shell (T) triples correction on gfx950: device kernels and their thin C ABI.
//
// Formulation (closed form of what src/triples-ijk/ijk_tuple.f90:56-84 builds with 12
// DGEMMs and trant3 steps):
//
//   W_ijk(a,b,c) = sum over the 6 simultaneous permutations s of the pairs (i,a),(j,b),(k,c)
//                  of  sum_K AX[x](alpha,beta,K) * BX[y,z](gamma,K)
//   AX[x](alpha,beta,K) = VVVO(K,beta,alpha,x)      K <  nu
//                       = T2(alpha,beta,K-nu,x)     nu <= K < nu+no
//   BX[y,z](gamma,K)    = T2(gamma,K,y,z)           K <  nu
//                       = -VM(K-nu,gamma,z,y)       nu <= K < nu+no
//
// with (x,alpha),(y,beta),(z,gamma) the permuted pairs.  Particle (VVVO*T2, K=nu) and
// hole (T2*VM, K=no) contractions are one GEMM over K = nu+no.
//
// Work item = (occupied tuple, virtual block orbit).  The virtual cube is cut into
// 16^3 blocks; a work item owns the blocks (A<=B<=C) and accumulates, for the six
// orderings P of its occupied tuple, w_P(a,b,c) = W_P(a,b,c) over (a,b,c) in A x B x C.
// Because W_{pi(P)}(pi(abc)) = W_P(abc), these six accumulators hold the values of
// W_ijk at all six positional permutations of (a,b,c), in the SAME register slot, so
// the energy and v1 need no transposition of W at all:
//   sum_s f(s(abc)) = Q(w) / D,  Q = 6 sum w_P^2 + 2 E^2 + 2 O^2 - 8 E O,
//   E = w over even orderings, O = w over odd orderings,
// where f is the summand of t1wt3_ijk's E[T] loop (src/triples-ijk/t1wt3.f90:143-159).
// The index permutations (trant3, src/common/common_cc.f90:62-333) become the choice
// of which LDS-staged operand panel a fragment is read from and in which orientation.
//
// FP64 MFMA: v_mfma_f64_16x16x4_f64.  Per 16x16 output tile, lane l supplies
// A[l%16][l/16], B[l/16][l%16] and holds D[l/16 + 4v][l%16], v=0..3 (checked on the device).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

typedef double d4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TB = 16;                 // virtual block edge
constexpr int NTHR = 256;              // 4 waves
constexpr int KS = 4;                  // K per MFMA step
constexpr int PSTR = 17;               // padded alpha stride of an LDS operand panel
constexpr int PANEL = KS * TB * PSTR;  // 1088 doubles
constexpr int BXCH = KS * TB;          // 64 doubles
constexpr int STAGE = 6 * PANEL + 6 * BXCH;
constexpr int TSTR_B = 17;             // transposition scratch strides
constexpr int TSTR_A = TB * TSTR_B;
constexpr int TCUBE = TB * TSTR_A;     // 4352 doubles
constexpr int VSTR = 17;
constexpr int VBLK = TB * VSTR;        // 272 doubles per staged VOE sub-block
constexpr int LDS_DOUBLES = (2 * STAGE > 3 * TCUBE) ? ((2 * STAGE > 36 * VBLK) ? 2 * STAGE : 36 * VBLK)
                                                    : ((3 * TCUBE > 36 * VBLK) ? 3 * TCUBE : 36 * VBLK);

// orderings of the occupied tuple positions (0,1,2); parity even for 0,3,4
__device__ constexpr int PERM[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

__host__ __device__ constexpr int pair_index(int y, int z) { return y * 2 + (z > y ? z - 1 : z); }

__host__ __device__ constexpr int perm_index(int a, int b, int c)
{
    return a == 0 ? (b == 1 ? 0 : 1) : a == 1 ? (b == 0 ? 2 : 3) : (b == 0 ? 4 : 5);
}

// Term tables per phase (0: gamma=b, 1: gamma=c, 2: gamma=a) and accumulator p:
// panel slot (orientation*3 + tuple position of x) and BX slot (ordered position pair).
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
    // position map m -> PERM[s][PERM[t][m]]
    int pi[3] = {PERM[s][PERM[t][0]], PERM[s][PERM[t][1]], PERM[s][PERM[t][2]]};
    return wsrc(p, pi);
}

template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

}  // namespace

struct TriplesArgs {
    const double *ax;      // [no][nb][nb][kp/4][4][16][16]
    const double *bx;      // [no][no][kp/4][4][nup]
    const double *voe;     // [no][no][nup][nup], voe[q][r][x][y] = VOE(x,y,q,r)
    const double *eh;      // [no]
    const double *ep;      // [nup], padded entries finite
    const int *tuples;     // [ntup][4]: i, j, k, 1 distinct / 2 degenerate
    const int *blocks;     // [nblk][4]: A, B, C, multiplicity code 1 / 2 / 6
    double *e_part;        // [nitems]
    double *v1_part;       // [nitems][3][16][3]
    double *wout;          // W_ijk cube (nu^3, column-major) of tuple 0, or null
    int no, nu, nup, nb, kp, ntup, nblk;
    long long nitems;
};

// ---------------------------------------------------------------------------------
// packing kernels (run once per problem)
// ---------------------------------------------------------------------------------
__global__ void pack_ax_kernel(const double *__restrict__ vvvo, const double *__restrict__ t2,
                               double *__restrict__ ax, int no, int nu, int nb, int kp)
{
    const long long nkb = kp / 4;
    const long long total = (long long)no * nb * nb * nkb * 1024;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx;
        const int beta = r & 15; r >>= 4;
        const int alpha = r & 15; r >>= 4;
        const int k4 = r & 3; r >>= 2;
        const long long kb = r % nkb; r /= nkb;
        const int q = r % nb; r /= nb;
        const int pblk = r % nb; r /= nb;
        const int x = (int)r;
        const int ag = pblk * TB + alpha, bg = q * TB + beta;
        const long long K = kb * 4 + k4;
        double v = 0.0;
        if (ag < nu && bg < nu) {
            if (K < nu)
                v = vvvo[K + (long long)nu * (bg + (long long)nu * (ag + (long long)nu * x))];
            else if (K < nu + no)
                v = t2[ag + (long long)nu * (bg + (long long)nu * ((K - nu) + (long long)no * x))];
        }
        ax[idx] = v;
    }
}

__global__ void pack_bx_kernel(const double *__restrict__ t2, const double *__restrict__ vm,
                               double *__restrict__ bx, int no, int nu, int nup, int kp)
{
    const long long total = (long long)no * no * kp * nup;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx;
        const int g = r % nup; r /= nup;
        const int K = r % kp; r /= kp;
        const int z = r % no; r /= no;
        const int y = (int)r;
        double v = 0.0;
        if (g < nu) {
            if (K < nu)
                v = t2[g + (long long)nu * (K + (long long)nu * (y + (long long)no * z))];
            else if (K < nu + no)
                v = -vm[(K - nu) + (long long)no * (g + (long long)nu * (z + (long long)no * y))];
        }
        bx[idx] = v;
    }
}

__global__ void pack_voe_kernel(const double *__restrict__ voe, double *__restrict__ out, int no,
                                int nu, int nup)
{
    const long long total = (long long)no * no * nup * nup;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx;
        const int y = r % nup; r /= nup;
        const int x = r % nup; r /= nup;
        const int rr = r % no; r /= no;
        const int q = (int)r;
        double v = 0.0;
        if (x < nu && y < nu)
            v = voe[x + (long long)nu * (y + (long long)nu * (q + (long long)no * rr))];
        out[idx] = v;
    }
}

// Distributed (nutr, nu*no) layout -> VVVO(nu,nu,nu,no): column c + nu*x holds the
// lower-packed symmetric (p,q) pair block, row p*(p+1)/2 + q for q <= p (0-based),
// as DDCC_T_GETVE / CPYTSQ expand it (src/triples-ijk/common_functions.f:43-60,106-125).
__global__ void unpack_vvvo_kernel(const double *__restrict__ packed, double *__restrict__ full, int no, int nu)
{
    const long long nutr = (long long)nu * (nu + 1) / 2;
    const long long total = (long long)nu * nu * nu * no;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int p = idx % nu;
        const int q = (idx / nu) % nu;
        const long long col = idx / ((long long)nu * nu);   // c + nu*x
        const int hi = p > q ? p : q, lo = p > q ? q : p;
        full[idx] = packed[(long long)hi * (hi + 1) / 2 + lo + nutr * col];
    }
}

// ---------------------------------------------------------------------------------
// persistent (tuple, block-orbit) kernel, FP64 MFMA
// ---------------------------------------------------------------------------------
namespace {

// Stage one K-step (KS values of K) of a phase's 6 operand panels and 6 BX chunks.
struct StageRegs {
    double pa[6][4];
    double bq[4];
};

__device__ __forceinline__ void stage_load(const TriplesArgs &g, StageRegs &st, const int *xs,
                                           int r0o0, int r1o0, int r0o1, int r1o1, int gblk,
                                           const int *ys, const int *zs, int kb, int tid)
{
    const long long nkb = g.kp / 4;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int o = s / 3, xp = s % 3;
        const int r0 = o ? r0o1 : r0o0, r1 = o ? r1o1 : r1o0;
        const double *src = g.ax + ((((long long)xs[xp] * g.nb + r0) * g.nb + r1) * nkb + kb) * 1024 + tid * 4;
        const double2 v0 = *reinterpret_cast<const double2 *>(src);
        const double2 v1 = *reinterpret_cast<const double2 *>(src + 2);
        st.pa[s][0] = v0.x; st.pa[s][1] = v0.y; st.pa[s][2] = v1.x; st.pa[s][3] = v1.y;
    }
    if (tid < 96) {
        const int s = tid >> 4, k = (tid >> 2) & 3, m0 = (tid & 3) * 4;
        const double *src = g.bx + (((long long)(ys[s] * g.no + zs[s]) * nkb + kb) * 4 + k) * g.nup + gblk * TB + m0;
        const double2 v0 = *reinterpret_cast<const double2 *>(src);
        const double2 v1 = *reinterpret_cast<const double2 *>(src + 2);
        st.bq[0] = v0.x; st.bq[1] = v0.y; st.bq[2] = v1.x; st.bq[3] = v1.y;
    }
}

__device__ __forceinline__ void stage_store(double *buf, const StageRegs &st, int tid)
{
    const int e = tid * 4;
    const int k = e >> 8, al = (e >> 4) & 15, be = e & 15;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        double *d = buf + s * PANEL + k * (TB * PSTR) + al * PSTR + be;
        d[0] = st.pa[s][0]; d[1] = st.pa[s][1]; d[2] = st.pa[s][2]; d[3] = st.pa[s][3];
    }
    if (tid < 96) {
        const int s = tid >> 4, k2 = (tid >> 2) & 3, m0 = (tid & 3) * 4;
        double *d = buf + 6 * PANEL + s * BXCH + k2 * TB + m0;
        d[0] = st.bq[0]; d[1] = st.bq[1]; d[2] = st.bq[2]; d[3] = st.bq[3];
    }
}

// One phase of the K loop.  PHASE 0: gamma=b (batch c), 1: gamma=c (batch b), 2: gamma=a (batch b).
template <int PHASE>
__device__ __forceinline__ void run_phase(const TriplesArgs &g, double *lds, d4 (&acc)[6][4],
                                          const int *xs, const int *ys, const int *zs,
                                          int A, int B, int C, int tid, int wave, int lane)
{
    int r0o0, r1o0, r0o1, r1o1, gblk;
    if (PHASE == 0) { r0o0 = A; r1o0 = C; r0o1 = C; r1o1 = A; gblk = B; }
    else if (PHASE == 1) { r0o0 = A; r1o0 = B; r0o1 = B; r1o1 = A; gblk = C; }
    else { r0o0 = B; r1o0 = C; r0o1 = C; r1o1 = B; gblk = A; }
    // fragment orientation: "row" reads P[k][batch][m], "col" reads P[k][m][batch]
    constexpr bool row0 = (PHASE == 2);
    constexpr bool row1 = (PHASE != 2);

    const int nsteps = g.kp / KS;
    const int kq = lane >> 4, m = lane & 15;
    StageRegs st;
    stage_load(g, st, xs, r0o0, r1o0, r0o1, r1o1, gblk, ys, zs, 0, tid);
    stage_store(lds, st, tid);
    __syncthreads();
    for (int s = 0; s < nsteps; ++s) {
        double *cur = lds + (s & 1) * STAGE;
        if (s + 1 < nsteps)
            stage_load(g, st, xs, r0o0, r1o0, r0o1, r1o1, gblk, ys, zs, s + 1, tid);
        double xf[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) xf[q] = cur[6 * PANEL + q * BXCH + kq * TB + m];
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int bt = wave * 4 + bb;
            double pf[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const bool row = (q < 3) ? row0 : row1;
                const double *pn = cur + q * PANEL + kq * (TB * PSTR);
                pf[q] = row ? pn[bt * PSTR + m] : pn[m * PSTR + bt];
            }
            static_for<6>([&](auto P) {
                constexpr int p = decltype(P)::value;
                constexpr Term t = term(PHASE, p);
                if constexpr (PHASE == 2) {
                    acc[p][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(xf[t.b1], pf[t.s1], acc[p][bb], 0, 0, 0);
                    acc[p][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(xf[t.b2], pf[t.s2], acc[p][bb], 0, 0, 0);
                } else {
                    acc[p][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(pf[t.s1], xf[t.b1], acc[p][bb], 0, 0, 0);
                    acc[p][bb] = __builtin_amdgcn_mfma_f64_16x16x4f64(pf[t.s2], xf[t.b2], acc[p][bb], 0, 0, 0);
                }
            });
        }
        if (s + 1 < nsteps) stage_store(lds + ((s + 1) & 1) * STAGE, st, tid);
        __syncthreads();
    }
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

}  // namespace

__global__ void __launch_bounds__(NTHR, 1) triples_f64_kernel(TriplesArgs g)
{
    extern __shared__ double lds[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ double red[4][16 * 3 * 4 + 16 * 3 + 16 * 3 + 1];

    for (long long item = blockIdx.x; item < g.nitems; item += gridDim.x) {
        const int t = (int)(item / g.nblk), q = (int)(item % g.nblk);
        const int4 tp = reinterpret_cast<const int4 *>(g.tuples)[t];
        const int4 bp = reinterpret_cast<const int4 *>(g.blocks)[q];
        const int tpos[3] = {tp.x, tp.y, tp.z};
        const int A = bp.x, B = bp.y, C = bp.z;
        // BX slots: ordered pairs of tuple positions
        int ys[6], zs[6];
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int z = 0; z < 3; ++z)
                if (y != z) { ys[pair_index(y, z)] = tpos[y]; zs[pair_index(y, z)] = tpos[z]; }

        d4 acc[6][4];
#pragma unroll
        for (int p = 0; p < 6; ++p)
#pragma unroll
            for (int bb = 0; bb < 4; ++bb) acc[p][bb] = d4{0.0, 0.0, 0.0, 0.0};

        // phase gamma=b in layout (batch c; rows a, cols b)
        run_phase<0>(g, lds, acc, tpos, ys, zs, A, B, C, tid, wave, lane);

        // transpose to layout (batch b; rows a, cols c) through LDS, 3 accumulators at a time
#pragma unroll
        for (int half = 0; half < 2; ++half) {
#pragma unroll
            for (int pp = 0; pp < 3; ++pp)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int a = (lane >> 4) + 4 * v, b = lane & 15, c = wave * 4 + bb;
                        lds[pp * TCUBE + a * TSTR_A + b * TSTR_B + c] = acc[half * 3 + pp][bb][v];
                    }
            __syncthreads();
#pragma unroll
            for (int pp = 0; pp < 3; ++pp)
#pragma unroll
                for (int bb = 0; bb < 4; ++bb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int a = (lane >> 4) + 4 * v, b = wave * 4 + bb, c = lane & 15;
                        acc[half * 3 + pp][bb][v] = lds[pp * TCUBE + a * TSTR_A + b * TSTR_B + c];
                    }
            __syncthreads();
        }

        run_phase<1>(g, lds, acc, tpos, ys, zs, A, B, C, tid, wave, lane);
        run_phase<2>(g, lds, acc, tpos, ys, zs, A, B, C, tid, wave, lane);

        // ---- stage the 36 VOE sub-blocks: [ordered tuple-position pair][ordered block pair]
        const int blk3[3] = {A, B, C};
        for (int idx = tid; idx < 36 * 16 * 4; idx += NTHR) {
            const int quad = idx & 3, row = (idx >> 2) & 15, sb = idx >> 6;
            const int qr = sb / 6, bpair = sb % 6;
            int y = 0, z = 0, ry = 0, rz = 0;
#pragma unroll
            for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                for (int zz = 0; zz < 3; ++zz)
                    if (yy != zz) {
                        if (pair_index(yy, zz) == qr) { y = yy; z = zz; }
                        if (pair_index(yy, zz) == bpair) { ry = yy; rz = zz; }
                    }
            const double *src = g.voe + (((long long)tpos[y] * g.no + tpos[z]) * g.nup + blk3[ry] * TB + row) * g.nup
                                + blk3[rz] * TB + quad * 4;
            double *d = lds + sb * VBLK + row * VSTR + quad * 4;
            d[0] = src[0]; d[1] = src[1]; d[2] = src[2]; d[3] = src[3];
        }
        __syncthreads();

        // ---- epilogue: energy, v1 partials, optional W store
        const double dijk = g.eh[tp.x] + g.eh[tp.y] + g.eh[tp.z];
        const double wgt = (tp.w == 1 ? 1.0 : 0.5) / (double)bp.w;
        double esum = 0.0;
        double va[4][3], vb[4][3], vc[3];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int p = 0; p < 3; ++p) { va[i][p] = 0.0; vb[i][p] = 0.0; }
        vc[0] = vc[1] = vc[2] = 0.0;
        const int cl = lane & 15;
        const int cg = C * TB + cl;
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) {
            const int bl = wave * 4 + bb;
            const int bg = B * TB + bl;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int al = (lane >> 4) + 4 * v;
                const int ag = A * TB + al;
                if (ag >= g.nu || bg >= g.nu || cg >= g.nu) continue;
                double w[6];
#pragma unroll
                for (int p = 0; p < 6; ++p) w[p] = acc[p][bb][v];
                const double inv = 1.0 / (dijk - g.ep[ag] - g.ep[bg] - g.ep[cg]);
                const double ev = w[0] + w[3] + w[4], od = w[1] + w[2] + w[5];
                const double sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3] + w[4] * w[4] + w[5] * w[5];
                esum += (6.0 * sq + 2.0 * ev * ev + 2.0 * od * od - 8.0 * ev * od) * inv;
                const int el[3] = {al, bl, cl};
                const double winv = inv;
                static_for<6>([&](auto S) {
                    constexpr int s = decltype(S)::value;
                    constexpr int r0 = PERM[s][0], r1 = PERM[s][1], r2 = PERM[s][2];
                    constexpr int bpair = pair_index(r1, r2);
                    static_for<6>([&](auto PP) {
                        constexpr int p = decltype(PP)::value;
                        constexpr int q0 = wsrc_comp(p, s, 0);   // sigma
                        constexpr int q1 = wsrc_comp(p, s, 2);   // sigma o (1,0,2)
                        constexpr int q2 = wsrc_comp(p, s, 1);   // sigma o (0,2,1)
                        constexpr int q3 = wsrc_comp(p, s, 3);   // sigma o (1,2,0)
                        constexpr int qr = pair_index(PERM[p][1], PERM[p][2]);
                        const double z = 2.0 * w[q0] - 2.0 * w[q1] - w[q2] + w[q3];
                        const double vv = lds[(qr * 6 + bpair) * VBLK + el[r1] * VSTR + el[r2]];
                        const double c = z * vv * winv;
                        if constexpr (r0 == 0) va[v][PERM[p][0]] += c;
                        else if constexpr (r0 == 1) vb[bb][PERM[p][0]] += c;
                        else vc[PERM[p][0]] += c;
                    });
                });
                if (g.wout != nullptr && t == 0) {
                    const int xg[3] = {ag, bg, cg};
                    static_for<6>([&](auto S) {
                        constexpr int s = decltype(S)::value;
                        constexpr int q0 = wsrc_comp(0, s, 0);
                        g.wout[xg[PERM[s][0]] + (long long)g.nu * (xg[PERM[s][1]] + (long long)g.nu * xg[PERM[s][2]])] = w[q0];
                    });
                }
            }
        }
        __syncthreads();   // VOE staging no longer needed; reuse red[]

        // ---- deterministic reductions
        double *rw = red[wave];
        // role a: rows (lane>>4)+4v, reduce over the 16 lanes of a group, then waves
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const double s = group16_sum(va[v][p]);
                if ((lane & 15) == 0) rw[((lane >> 4) + 4 * v) * 3 + p] = s;
            }
        // role b: rows wave*4+bb, whole-wave sums
#pragma unroll
        for (int bb = 0; bb < 4; ++bb)
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const double s = wave_sum(vb[bb][p]);
                if (lane == 0) rw[48 + bb * 3 + p] = s;
            }
        // role c: rows lane&15, sum over the 4 lane groups, then waves
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            double s = vc[p];
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (lane < 16) rw[60 + lane * 3 + p] = s;
        }
        {
            const double s = wave_sum(esum);
            if (lane == 0) rw[108] = s;
        }
        __syncthreads();
        double *vp = g.v1_part + item * 144;
        if (tid < 48) {            // role a
            const double s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
            vp[tid] = s * wgt;
        } else if (tid < 96) {     // role b: row = tid-48 -> wave (row/4), bb (row%4)
            const int r = tid - 48, row = r / 3, p = r % 3;
            vp[48 + r] = red[row >> 2][48 + (row & 3) * 3 + p] * wgt;
        } else if (tid < 144) {    // role c
            const int r = tid - 96;
            const double s = red[0][60 + r] + red[1][60 + r] + red[2][60 + r] + red[3][60 + r];
            vp[96 + r] = s * wgt;
        } else if (tid == 144) {
            g.e_part[item] = (red[0][108] + red[1][108] + red[2][108] + red[3][108]) * wgt;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// finalize: deterministic reductions
// ---------------------------------------------------------------------------------
// v1t[t][a][p] = sum over the blocks of tuple t whose role-r block contains a
__global__ void reduce_v1_tuple_kernel(const double *__restrict__ v1_part, const int *__restrict__ blocks,
                                       double *__restrict__ v1t, int ntup, int nblk, int nu)
{
    const long long total = (long long)ntup * nu * 3;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int p = idx % 3;
        const int a = (idx / 3) % nu;
        const long long t = idx / (3LL * nu);
        const int ab = a / TB, al = a % TB;
        double s = 0.0;
        for (int q = 0; q < nblk; ++q) {
            const int4 bp = reinterpret_cast<const int4 *>(blocks)[q];
            const double *vp = v1_part + (t * nblk + q) * 144;
            if (bp.x == ab) s += vp[al * 3 + p];
            if (bp.y == ab) s += vp[48 + al * 3 + p];
            if (bp.z == ab) s += vp[96 + al * 3 + p];
        }
        v1t[idx] = s;
    }
}

// v1(a,o) (column-major nu x no) = sum over tuples and positions p with tuple[p] == o
__global__ void reduce_v1_kernel(const double *__restrict__ v1t, const int *__restrict__ tuples,
                                 double *__restrict__ v1, int ntup, int nu, int no)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nu * no) return;
    const int a = idx % nu, o = idx / nu;
    double s = 0.0;
    for (int t = 0; t < ntup; ++t) {
        const int4 tp = reinterpret_cast<const int4 *>(tuples)[t];
        const double *r = v1t + ((long long)t * nu + a) * 3;
        if (tp.x == o) s += r[0];
        if (tp.y == o) s += r[1];
        if (tp.z == o) s += r[2];
    }
    v1[idx] = s;
}

// out[0] = etd = sum e_part (compensated, fixed order); out[1] = ets = 2 sum T1(o,a) v1(a,o)
__global__ void __launch_bounds__(1024) finalize_energy_kernel(const double *__restrict__ e_part, long long n,
                                                               const double *__restrict__ t1,
                                                               const double *__restrict__ v1, int no, int nu,
                                                               double *__restrict__ out)
{
    __shared__ double sh[2][1024];
    const int tid = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (long long i = tid; i < n; i += 1024) {
        const double y = e_part[i];
        const double tt = s + y;
        c += (fabs(s) >= fabs(y)) ? ((s - tt) + y) : ((y - tt) + s);
        s = tt;
    }
    double s2 = 0.0, c2 = 0.0;
    for (int i = tid; i < no * nu; i += 1024) {
        const int a = i % nu, o = i / nu;
        const double y = t1[o + no * a] * v1[i];
        const double tt = s2 + y;
        c2 += (fabs(s2) >= fabs(y)) ? ((s2 - tt) + y) : ((y - tt) + s2);
        s2 = tt;
    }
    sh[0][tid] = s + c;
    sh[1][tid] = s2 + c2;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if (tid < w) {
            sh[0][tid] += sh[0][tid + w];
            sh[1][tid] += sh[1][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        out[0] = sh[0][0];
        out[1] = 2.0 * sh[1][0];
    }
}

// ---------------------------------------------------------------------------------
// thin C ABI used by the C host layer
// ---------------------------------------------------------------------------------
extern "C" {

int ccx_hip_device_count(int *n) { return hipGetDeviceCount(n) == hipSuccess ? 0 : 1; }

int ccx_hip_malloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return 0;
    return hipMalloc(p, bytes) == hipSuccess ? 0 : 1;
}

int ccx_hip_free(void *p) { return (p == nullptr || hipFree(p) == hipSuccess) ? 0 : 1; }

int ccx_hip_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream) != hipSuccess) return 1;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess) return 1;
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_sync(void *stream) { return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : 1; }

int ccx_hip_cu_count(int *n)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    return hipDeviceGetAttribute(n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess ? 0 : 1;
}

const char *ccx_hip_last_error(void) { return hipGetErrorString(hipGetLastError()); }

static int grid_for(long long total)
{
    long long g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

int ccx_hip_pack(const double *vvvo, const double *t2, const double *vm, const double *voe, double *ax,
                 double *bx, double *voep, int no, int nu, int nup, int kp, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int nb = nup / TB;
    pack_ax_kernel<<<grid_for((long long)no * nb * nb * kp * 256), 256, 0, s>>>(vvvo, t2, ax, no, nu, nb, kp);
    pack_bx_kernel<<<grid_for((long long)no * no * kp * nup), 256, 0, s>>>(t2, vm, bx, no, nu, nup, kp);
    pack_voe_kernel<<<grid_for((long long)no * no * nup * nup), 256, 0, s>>>(voe, voep, no, nu, nup);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_triples_f64(const double *ax, const double *bx, const double *voep, const double *eh,
                        const double *ep, const int *tuples, const int *blocks, double *e_part, double *v1_part,
                        double *wout, int no, int nu, int nup, int kp, int ntup, int nblk, int grid,
                        void *stream)
{
    static bool attr_set = false;
    const size_t lds_bytes = sizeof(double) * LDS_DOUBLES;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)triples_f64_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_bytes) != hipSuccess)
            return 2;
        attr_set = true;
    }
    TriplesArgs g;
    g.ax = ax; g.bx = bx; g.voe = voep; g.eh = eh; g.ep = ep; g.tuples = tuples; g.blocks = blocks;
    g.e_part = e_part; g.v1_part = v1_part; g.wout = wout;
    g.no = no; g.nu = nu; g.nup = nup; g.nb = nup / TB; g.kp = kp; g.ntup = ntup; g.nblk = nblk;
    g.nitems = (long long)ntup * nblk;
    if (g.nitems == 0) return 0;
    if (grid <= 0 || grid > g.nitems) grid = (int)(g.nitems < 256 ? g.nitems : 256);
    triples_f64_kernel<<<grid, NTHR, lds_bytes, (hipStream_t)stream>>>(g);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_finalize(const double *e_part, const double *v1_part, const int *tuples, const int *blocks,
                     const double *t1, double *v1t, double *v1, double *out, int no, int nu, int ntup, int nblk,
                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (ntup > 0) {
        reduce_v1_tuple_kernel<<<grid_for((long long)ntup * nu * 3), 256, 0, s>>>(v1_part, blocks, v1t, ntup,
                                                                                  nblk, nu);
        reduce_v1_kernel<<<(nu * no + 255) / 256, 256, 0, s>>>(v1t, tuples, v1, ntup, nu, no);
    } else {
        if (hipMemsetAsync(v1, 0, sizeof(double) * nu * no, s) != hipSuccess) return 1;
    }
    finalize_energy_kernel<<<1, 1024, 0, s>>>(e_part, (long long)ntup * nblk, t1, v1, no, nu, out);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_lds_bytes(void) { return (int)(sizeof(double) * LDS_DOUBLES); }

int ccx_hip_unpack_vvvo(const double *packed, double *full, int no, int nu, void *stream)
{
    unpack_vvvo_kernel<<<grid_for((long long)nu * nu * nu * no), 256, 0, (hipStream_t)stream>>>(packed, full, no,
                                                                                                 nu);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_event_create(void **ev)
{
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return 1;
    *ev = (void *)e;
    return 0;
}

int ccx_hip_event_destroy(void *ev) { return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? 0 : 1; }

int ccx_hip_event_record(void *ev, void *stream)
{
    return hipEventRecord((hipEvent_t)ev, (hipStream_t)stream) == hipSuccess ? 0 : 1;
}

int ccx_hip_event_ms(void *start, void *stop, float *ms)
{
    if (hipEventSynchronize((hipEvent_t)stop) != hipSuccess) return 1;
    return hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop) == hipSuccess ? 0 : 1;
}

}  // extern "C"
