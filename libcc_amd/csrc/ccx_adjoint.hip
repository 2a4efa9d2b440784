// Adjoint of the closed-shell (T) energy on gfx950: dE/d(T1, T2, VM, VOE, VVVO).
//
// E = E[T] + E_ST with (per tuple t of weight wt, canonical cube W_t)
//   E[T] = sum_t wt sum_abc X_t,   X = W (8W - 4(W_acb+W_cba+W_bac) + 2(W_bca+W_cab)) / D
//   E_ST = 2 sum T1 . v1,          v1(a,p) = sum Z_pqr(a,b,c) VOE(b,c,q,r)
// (src/triples-ijk/t1wt3.f90:53-102, src/triples-ijk/triples.F90:310-313).  The tuple
// loop is re-run: W is recomputed per work item exactly as in the forward kernel and
// never stored across tuples.  Per batch of tuples:
//
//  1. adj_item_kernel  (tuple, block orbit) items: W in registers, then in the epilogue
//     Wbar_t(x) = dE/dW_t(x) at the six positional permutations of every point,
//       Wbar = wt [ (12 W_s + 4 E' - 8 O') / D  +  2/D sum_{p,mu} c_mu T1.VOE ],
//     written to the batch's blocked Wbar cube; the T1.Z contraction for dVOE reduced
//     over the item's a-block into per-(tuple, ordering, block triple) slots; energy and
//     v1 partials as the forward kernel.
//  2. gax_kernel       gAX[x](ab,K) += sum_g Wbar_sigma(a,b,g) BX[y,z](g,K)
//  3. gbx_kernel       gBX[y,z](g,K) += sum_ab Wbar_sigma(a,b,g) AX[x](a,b,K)
//     (sigma over the six simultaneous permutations that build W from AX and BX).
//  4. dvoe_reduce_kernel  slots -> dVOE.
// After the last batch unpack_grad_kernel maps gAX/gBX back to VVVO, T2 and VM.
// Every accumulation is in a fixed order: the result is bitwise reproducible.
#include "ccx_item.h"

namespace {

constexpr int WB_STR_A = 273;                      // padded strides of a 16^3 Wbar block in LDS
constexpr int WB_STR_B = 17;
constexpr int WB_BLK = TB * WB_STR_A;              // 4368 doubles
constexpr int ADJ_V = 0;                           // LDS regions of the item epilogue (doubles)
constexpr int ADJ_D = 36 * VBLK;                   // 36 staged VOE sub-blocks
constexpr int ADJ_T = ADJ_D + 36 * 256;            // 36 dVOE partial sub-blocks
constexpr int ADJ_END = ADJ_T + 144;               // T1 rows of the tuple over the item's blocks
constexpr int ADJ_LDS = ADJ_END > 3 * CUBE ? ADJ_END : 3 * CUBE;

// Z terms of the ets adjoint: Z(x) = sum_m c_m W(mu_m x) / D with
// mu = id, (1,0,2), (0,2,1), (1,2,0) and c = 2, -2, -1, 1.
__device__ constexpr int ZMU[4][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {1, 2, 0}};
__device__ constexpr double ZC[4] = {2.0, -2.0, -1.0, 1.0};

__host__ __device__ constexpr int inv_perm_at(const int *p, int n) { return p[0] == n ? 0 : p[1] == n ? 1 : 2; }

// Placement index of the T1.VOE product that Z term m of ordering p contributes to
// Wbar at placement s: n = PERM[s] o lam^-1 with lam(n') = mu_m(p^-1(n')).
__host__ __device__ constexpr int sn_index(int s, int p, int m)
{
    int lam[3] = {0, 0, 0};
    for (int n = 0; n < 3; ++n) lam[n] = ZMU[m][inv_perm_at(PERM[p], n)];
    int li[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) li[k] = inv_perm_at(lam, k);
    return perm_index(PERM[s][li[0]], PERM[s][li[1]], PERM[s][li[2]]);
}

// index of PERM[s]^-1
__host__ __device__ constexpr int perm_inverse(int s)
{
    return perm_index(inv_perm_at(PERM[s], 0), inv_perm_at(PERM[s], 1), inv_perm_at(PERM[s], 2));
}

// sigma-terms: W_t(e) += sum_K AX[t[xp]](e_u, e_v, K) BX[t[yp], t[zp]](e_wf, K)
struct Sigma { int xp, u, v, yp, zp, wf; };
__device__ constexpr Sigma SIG[6] = {{0, 0, 1, 1, 2, 2}, {1, 1, 0, 0, 2, 2}, {2, 2, 1, 1, 0, 0},
                                     {0, 0, 2, 2, 1, 1}, {1, 1, 2, 2, 0, 0}, {2, 2, 0, 0, 1, 1}};

}  // namespace

struct AdjArgs {
    TriplesArgs f;          // forward operands; e_part / v1_part indexed by global item
    const double *t1;       // T1(no, nu), column-major
    double *wbar;           // [tcount][nb][nb][nb][16][16][16]
    double *slots;          // [tcount][6][nb][nb][nb][16][16]  dVOE partials
    double *wstage;         // [grid][6][BPW][4][NTHR]: the item's W, one slab per work group
    int t0, tcount;         // tuples of this batch
};

namespace {

// Epilogue of one adjoint work item; acc holds w_p in layout L.
// The item's W is read back per point from the work group's staging slab ws (written by
// the same thread), so that no accumulator registers stay live through the epilogue.
__device__ __forceinline__ double staged_w(const double *ws, int p, int bb, int v, int tid)
{
    return ws[((p * BPW + bb) * 4 + v) * NTHR + tid];
}

__device__ void adj_epilogue(const AdjArgs &ga, double *lds, double (*red)[112], const ItemCtx &it,
                             long long item, int tl, int tkind, int mult, bool swap_bc, const double *ws, int wave,
                             int lane, int tid)
{
    const TriplesArgs &g = ga.f;
    const int nb = g.nb;
    double *V = lds + ADJ_V, *D = lds + ADJ_D, *T = lds + ADJ_T;
    // ---- stage: 36 VOE sub-blocks (as the forward epilogue), T1 rows, zero dVOE partials
    for (int idx = tid; idx < 36 * 16 * 4; idx += NTHR) {
        const int quad = idx & 3, row = (idx >> 2) & 15, sb = idx >> 6;
        const int qr = sb / 6, bpair = sb % 6;
        const int y = qr >> 1, z = (qr & 1) ? (y == 2 ? 1 : 2) : (y == 0 ? 1 : 0);
        const int ry = bpair >> 1, rz = (bpair & 1) ? (ry == 2 ? 1 : 2) : (ry == 0 ? 1 : 0);
        const double *src = g.voe + (((long long)it.tpos[y] * g.no + it.tpos[z]) * g.nup + it.blk[ry] * TB + row) * g.nup
                            + it.blk[rz] * TB + quad * 4;
        double *d = V + sb * VBLK + row * VSTR + quad * 4;
        d[0] = src[0]; d[1] = src[1]; d[2] = src[2]; d[3] = src[3];
    }
    if (tid < 144) {          // T[(y*3+n)*16+l] = T1(tuple[y], blk[n]*16+l)
        const int l = tid & 15, yn = tid >> 4, y = yn / 3, n = yn % 3;
        const int a = it.blk[n] * TB + l;
        T[tid] = a < g.nu ? ga.t1[it.tpos[y] + (long long)g.no * a] : 0.0;
    }
    for (int idx = tid; idx < 36 * 256; idx += NTHR) D[idx] = 0.0;
    for (int idx = lane; idx < 112; idx += 64) red[wave][idx] = 0.0;
    __syncthreads();

    const double dijk = g.eh[it.tpos[0]] + g.eh[it.tpos[1]] + g.eh[it.tpos[2]];
    const double wt = tkind == 1 ? 1.0 : 0.5;
    const double wgt = wt / (double)mult;
    const int cl = lane & 15;
    const int A = it.blk[0], B = it.blk[1], C = it.blk[2];
    const int cg = C * TB + cl;
    double *wb = ga.wbar + (long long)tl * nb * nb * nb * 4096;
    double *rw = red[wave];
    double esum = 0.0;

    // The three passes run over this thread's 8 points (rows a = row(lane,v), batches
    // b = wave*BPW+bb, column c = lane&15) with a run-time loop over v, so only one row
    // of points is live at a time.
    // pass 1: energy and v1 partials (the forward epilogue), reduced per point:
    //   role a (rows al): over the 16 lanes of a row group  -> rw[al*3 + p]
    //   role b (rows bl): over the wave                     -> rw[48 + bb*3 + p]
    //   role c (rows cl): over the four row groups          -> rw[54 + cl*3 + p]
#pragma unroll 1
    for (int v = 0; v < 4; ++v) {
        const int al = Mfma<double>::row(lane, v);
        const int ag = A * TB + al;
        static_for<BPW>([&](auto BB) __attribute__((always_inline)) {
            constexpr int bb = decltype(BB)::value;
            const int bl = wave * BPW + bb;
            const int bg = B * TB + bl;
            double w[6];
#pragma unroll
            for (int p = 0; p < 6; ++p) w[p] = staged_w(ws, p, bb, v, tid);
            const double inv = 1.0 / (dijk - g.ep[ag] - g.ep[bg] - g.ep[cg]);
            if (ag < g.nu && bg < g.nu && cg < g.nu) {
                const double ev = w[0] + w[3] + w[4], od = w[1] + w[2] + w[5];
                const double sq = w[0] * w[0] + w[1] * w[1] + w[2] * w[2] + w[3] * w[3] + w[4] * w[4] + w[5] * w[5];
                esum += (6.0 * sq + 2.0 * ev * ev + 2.0 * od * od - 8.0 * ev * od) * inv;
            }
            const int el[3] = {al, bl, cl};
            double ra[3] = {0, 0, 0}, rb[3] = {0, 0, 0}, rc[3] = {0, 0, 0};
            static_for<6>([&](auto S) __attribute__((always_inline)) {
                constexpr int s = decltype(S)::value;
                constexpr int r0 = PERM[s][0], r1 = PERM[s][1], r2 = PERM[s][2];
                constexpr int bpair = pair_index(r1, r2);
                static_for<6>([&](auto PP) __attribute__((always_inline)) {
                    constexpr int p = decltype(PP)::value;
                    constexpr int q0 = wsrc_comp(p, s, 0), q1 = wsrc_comp(p, s, 2);
                    constexpr int q2 = wsrc_comp(p, s, 1), q3 = wsrc_comp(p, s, 3);
                    constexpr int qr = pair_index(PERM[p][1], PERM[p][2]);
                    const double z = 2.0 * w[q0] - 2.0 * w[q1] - w[q2] + w[q3];
                    const double cv = z * inv * V[(qr * 6 + bpair) * VBLK + el[r1] * VSTR + el[r2]];
                    if constexpr (r0 == 0) ra[PERM[p][0]] += cv;
                    else if constexpr (r0 == 1) rb[PERM[p][0]] += cv;
                    else rc[PERM[p][0]] += cv;
                });
            });
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const double sa = group16_sum(ra[p]);
                if (cl == 0) rw[al * 3 + p] += sa;
                const double sb = wave_sum(rb[p]);
                if (lane == 0) rw[48 + bb * 3 + p] += sb;
                double sc = rc[p];
                sc += __shfl_xor(sc, 16, 64);
                sc += __shfl_xor(sc, 32, 64);
                if (lane < 16) rw[54 + cl * 3 + p] += sc;
            }
        });
    }
    {
        const double s = wave_sum(esum);
        if (lane == 0) rw[102] = s;
    }
    __syncthreads();
    double *vp = g.v1_part + item * 144;
    const int slot_b = swap_bc ? 96 : 48, slot_c = swap_bc ? 48 : 96;   // table role order
    if (tid < 48) {
        double s = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < NWAVE; ++w2) s += red[w2][tid];
        vp[tid] = s * wgt;
    } else if (tid < 96) {
        const int r = tid - 48, row = r / 3, p = r % 3;
        vp[slot_b + r] = red[row / BPW][48 + (row % BPW) * 3 + p] * wgt;
    } else if (tid < 144) {
        const int r = tid - 96;
        double s = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < NWAVE; ++w2) s += red[w2][54 + r];
        vp[slot_c + r] = s * wgt;
    } else if (tid == 144) {
        double s = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < NWAVE; ++w2) s += red[w2][102];
        g.e_part[item] = s * wgt;
    }

    // pass 2: Wbar at the six placements x^s = (x[PERM[s][0]], x[PERM[s][1]], x[PERM[s][2]])
#pragma unroll 1
    for (int v = 0; v < 4; ++v) {
        const int al = Mfma<double>::row(lane, v);
        const int ag = A * TB + al;
        static_for<BPW>([&](auto BB) __attribute__((always_inline)) {
            constexpr int bb = decltype(BB)::value;
            const int bl = wave * BPW + bb;
            const int bg = B * TB + bl;
            const int el[3] = {al, bl, cl};
            double wets[6] = {0, 0, 0, 0, 0, 0};   // T1.VOE part of Wbar
            static_for<6>([&](auto S) __attribute__((always_inline)) {
                constexpr int s = decltype(S)::value;
                constexpr int r0 = PERM[s][0], r1 = PERM[s][1], r2 = PERM[s][2];
                constexpr int bpair = pair_index(r1, r2);
                static_for<6>([&](auto PP) __attribute__((always_inline)) {
                    constexpr int p = decltype(PP)::value;
                    constexpr int qr = pair_index(PERM[p][1], PERM[p][2]);
                    const double pr = T[(PERM[p][0] * 3 + r0) * 16 + el[r0]] *
                                      V[(qr * 6 + bpair) * VBLK + el[r1] * VSTR + el[r2]];
                    static_for<6>([&](auto SS) __attribute__((always_inline)) {
                        constexpr int s2 = decltype(SS)::value;
                        static_for<4>([&](auto M) __attribute__((always_inline)) {
                            constexpr int m = decltype(M)::value;
                            if constexpr (sn_index(s2, p, m) == s) wets[s2] += ZC[m] * pr;
                        });
                    });
                });
            });
            double w[6];
#pragma unroll
            for (int p = 0; p < 6; ++p) w[p] = staged_w(ws, p, bb, v, tid);
            const double inv = 1.0 / (dijk - g.ep[ag] - g.ep[bg] - g.ep[cg]);
            const double ev = w[0] + w[3] + w[4], od = w[1] + w[2] + w[5];
            const bool diag = ag == bg && bg == cg;
            static_for<6>([&](auto S) __attribute__((always_inline)) {
                constexpr int s = decltype(S)::value;
                constexpr bool even = (s == 0 || s == 3 || s == 4);
                const double ws_ = w[perm_inverse(s)];
                const double same = even ? ev : od, other = even ? od : ev;
                const double val = diag ? 0.0 : wt * inv * ((12.0 * ws_ + 4.0 * same - 8.0 * other) + 2.0 * wets[s]);
                const int e0 = el[PERM[s][0]], e1 = el[PERM[s][1]], e2 = el[PERM[s][2]];
                const int X = it.blk[PERM[s][0]], Y = it.blk[PERM[s][1]], Z = it.blk[PERM[s][2]];
                // with equal blocks several placements cover the same block triple, from other
                // threads and with other rounding: only the first placement writes it
                bool first = true;
#pragma unroll
                for (int s2 = 0; s2 < s; ++s2)
                    if (it.blk[PERM[s2][0]] == X && it.blk[PERM[s2][1]] == Y && it.blk[PERM[s2][2]] == Z) first = false;
                if (first) wb[(((long long)X * nb + Y) * nb + Z) * 4096 + (e0 * 16 + e1) * 16 + e2] = val;
            });
        });
    }

    __syncthreads();   // every wave is done with the staged VOE blocks

    // pass 3 (after the Wbar pass, so that the staged VOE region is free for the b-axis
    // partials): dVOE partials, dVOE(x_r1, x_r2, t[P1], t[P2]) += Z inv T1(t[P0], x_r0), summed
    // over the item's block of the contracted axis r0
#pragma unroll 1
    for (int v = 0; v < 4; ++v) {
        const int al = Mfma<double>::row(lane, v);
        const int ag = A * TB + al;
        double dB[2][6];      // b-axis contracted: summed over batches, then waves
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int p = 0; p < 6; ++p) dB[s][p] = 0.0;
        static_for<BPW>([&](auto BB) __attribute__((always_inline)) {
            constexpr int bb = decltype(BB)::value;
            const int bl = wave * BPW + bb;
            const int bg = B * TB + bl;
            double w[6];
#pragma unroll
            for (int p = 0; p < 6; ++p) w[p] = staged_w(ws, p, bb, v, tid);
            const double inv = 1.0 / (dijk - g.ep[ag] - g.ep[bg] - g.ep[cg]);
            const int el[3] = {al, bl, cl};
            static_for<6>([&](auto S) __attribute__((always_inline)) {
                constexpr int s = decltype(S)::value;
                constexpr int r0 = PERM[s][0], r1 = PERM[s][1], r2 = PERM[s][2];
                static_for<6>([&](auto PP) __attribute__((always_inline)) {
                    constexpr int p = decltype(PP)::value;
                    constexpr int q0 = wsrc_comp(p, s, 0), q1 = wsrc_comp(p, s, 2);
                    constexpr int q2 = wsrc_comp(p, s, 1), q3 = wsrc_comp(p, s, 3);
                    const double z = 2.0 * w[q0] - 2.0 * w[q1] - w[q2] + w[q3];
                    const double dv = z * inv * T[(PERM[p][0] * 3 + r0) * 16 + el[r0]];
                    if constexpr (s < 2) {
                        // a-axis: sum over the four row groups, then over v by the same lane
                        // (s = 0 -> target (b, c); s = 1 -> (c, b))
                        double r = dv;
                        r += __shfl_xor(r, 16, 64);
                        r += __shfl_xor(r, 32, 64);
                        if (lane < 16) D[(s * 6 + p) * 256 + el[r1] * 16 + el[r2]] += r;
                    } else if constexpr (s < 4) {
                        dB[s - 2][p] += dv;
                    } else {
                        // c-axis: sum over the 16 lanes of equal row (s = 4 -> (a,b); 5 -> (b,a))
                        double r = dv;
                        r += __shfl_xor(r, 1, 64);
                        r += __shfl_xor(r, 2, 64);
                        r += __shfl_xor(r, 4, 64);
                        r += __shfl_xor(r, 8, 64);
                        if (cl == 0) D[(s * 6 + p) * 256 + el[r1] * 16 + el[r2]] = r;
                    }
                });
            });
        });
        // b-axis: each wave's partials to its own slot of PB (the staged VOE region, no longer
        // read), then the 12 x 64 targets of this row step summed over the waves in wave
        // order (deterministic)
        double *PB = V;   // [NWAVE][12][64]
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int p = 0; p < 6; ++p) PB[(wave * 12 + s * 6 + p) * 64 + lane] = dB[s][p];
        __syncthreads();
        for (int t = tid; t < 12 * 64; t += NTHR) {
            const int ln = t & 63, k = t >> 6, s = k / 6, p = k % 6;
            double sum = 0.0;
            for (int w2 = 0; w2 < NWAVE; ++w2) sum += PB[(w2 * 12 + k) * 64 + ln];
            // s = 2: (1,0,2) -> target (a, c); s = 3: (1,2,0) -> target (c, a)
            const int a_l = Mfma<double>::row(ln, v), c_l = ln & 15;
            const int idx = s == 0 ? a_l * 16 + c_l : c_l * 16 + a_l;
            D[((s + 2) * 6 + p) * 256 + idx] += sum;
        }
        __syncthreads();
    }

    // ---- dVOE partials -> slots[tl][p][X][Y][Z]; placements that land on the same block
    // triple (equal blocks) are summed here
    const double f2 = 2.0 * wgt;
    double *sl = ga.slots + (long long)tl * 6 * nb * nb * nb * 256;
    for (int idx = tid; idx < 36 * 256; idx += NTHR) {
        const int sp = idx >> 8, e = idx & 255;
        const int s = sp / 6, p = sp % 6;
        const int X = it.blk[PERM[s][1]], Y = it.blk[PERM[s][2]], Z = it.blk[PERM[s][0]];
        bool first = true;
        for (int s2 = 0; s2 < s; ++s2)
            if (it.blk[PERM[s2][1]] == X && it.blk[PERM[s2][2]] == Y && it.blk[PERM[s2][0]] == Z) first = false;
        if (!first) continue;
        double sum = 0.0;
        for (int s2 = s; s2 < 6; ++s2)
            if (it.blk[PERM[s2][1]] == X && it.blk[PERM[s2][2]] == Y && it.blk[PERM[s2][0]] == Z)
                sum += D[(s2 * 6 + p) * 256 + e];
        sl[((((long long)p * nb + X) * nb + Y) * nb + Z) * 256 + e] = sum * f2;
    }
    __syncthreads();
}

}  // namespace

// FK 1 / 2: W is recomputed by the FP32 or the split BF16x3 K sweep (FP32 accumulation,
// layout row 4 (lane >> 4) + v) and staged, as FP64, in the FP64 layout (row (lane >> 4) + 4 v)
// that the epilogue reads; everything after the staging is the FP64 path.  FK 0: FP64 sweep.
template <int FK>
__global__ void __launch_bounds__(NTHR, 1) adj_item_kernel(AdjArgs ga)
{
    extern __shared__ double lds[];
    __shared__ double red[NWAVE][112];
    __shared__ long long next_item;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const TriplesArgs &g = ga.f;
    const long long nlocal = (long long)ga.tcount * g.nblk;
    for (;;) {
        if (tid == 0) next_item = (long long)atomicAdd(g.counter, 1u);
        __syncthreads();
        const long long li = next_item;
        __syncthreads();
        if (li >= nlocal) break;
        const int tl = __builtin_amdgcn_readfirstlane((int)(li / g.nblk));
        const int q = __builtin_amdgcn_readfirstlane((int)(li % g.nblk));
        const int t = ga.t0 + tl;
        const long long item = (long long)t * g.nblk + q;
        int tkind, mult;
        ItemCtx it = make_item_ctx(g.tuples, g.blocks, t, q, tkind, mult);
        const bool swap_bc = ragged_block_to_b(g, it, mult);   // as the forward kernel does
        double *ws = ga.wstage + (long long)blockIdx.x * (6 * BPW * 4 * NTHR);
        if constexpr (FK != 0) {
            f4 acc[6][BPW];
            if constexpr (FK == 2) {
                f4 acc1[1][BPW];   // unused: the BF16 x 3 form has one accumulator
                static_assert(!pair_modes(2), "the forward kernel's modes");
                if (tkind == 2) item_w_split<M_IJJ, 2>(g, lds, it, acc, acc1, wave, lane);
                else if (tkind == 3) item_w_split<M_IIJ, 2>(g, lds, it, acc, acc1, wave, lane);
                else if (mult == 1) item_w_split<M_FULL, 2>(g, lds, it, acc, acc1, wave, lane);
                else if (mult == 6) item_w_split<M_AAA, 2>(g, lds, it, acc, acc1, wave, lane);
                else if (it.blk[0] == it.blk[1]) item_w_split<M_AAC, 2>(g, lds, it, acc, acc1, wave, lane);
                else item_w_split<M_ACC, 2>(g, lds, it, acc, acc1, wave, lane);
            } else {
                static_assert(!pair_modes(1), "the forward kernel's modes");
                if (tkind == 2) item_w_f32v<M_IJJ>(g, lds, it, acc, wave, lane);
                else if (tkind == 3) item_w_f32v<M_IIJ>(g, lds, it, acc, wave, lane);
                else if (mult == 1) item_w_f32v<M_FULL>(g, lds, it, acc, wave, lane);
                else if (mult == 6) item_w_f32v<M_AAA>(g, lds, it, acc, wave, lane);
                else if (it.blk[0] == it.blk[1]) item_w_f32v<M_AAC>(g, lds, it, acc, wave, lane);
                else item_w_f32v<M_ACC>(g, lds, it, acc, wave, lane);
            }
            // element (row 4 (lane >> 4) + v, column lane & 15) belongs to FP64-layout lane
            // 16 v + (lane & 15), slot lane >> 4 of the same wave
            const int dst = wave * 64 + (lane & 15);
#pragma unroll
            for (int p = 0; p < 6; ++p)
#pragma unroll
                for (int bb = 0; bb < BPW; ++bb)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        ws[((p * BPW + bb) * 4 + (lane >> 4)) * NTHR + dst + 16 * v] = (double)acc[p][bb][v] * g.wscale;
            __syncthreads();
        } else {
            d4 acc[6][BPW];
            static_assert(pair_modes(0), "the forward kernel's modes");
            dispatch_mode(tkind, mult, it, [&](auto M) __attribute__((always_inline)) {
                item_w<decltype(M)::value, double>(g, lds, it, acc, wave, lane);
            });
#pragma unroll
            for (int p = 0; p < 6; ++p)
#pragma unroll
                for (int bb = 0; bb < BPW; ++bb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) ws[((p * BPW + bb) * 4 + v) * NTHR + tid] = acc[p][bb][v];
        }
        adj_epilogue(ga, lds, red, it, item, tl, tkind, mult, swap_bc, ws, wave, lane, tid);
    }
}

// ---------------------------------------------------------------------------------
// back-contraction GEMMs (FP64 MFMA, LDS-staged, double buffered)
// ---------------------------------------------------------------------------------
namespace {
constexpr int GTHR = 512;                 // GEMM work group: 8 waves, 2 per SIMD
constexpr int GBPW = 2;                   // gax: rows b per wave
constexpr int GCH = 128;                  // K columns per work group (8 MFMA tiles)
constexpr int GCT = GCH / 16;
constexpr int BS_STR = 17;                // padded row of the staged BX / AX slices
}  // namespace

struct GemmArgs {
    const double *wbar;     // batch Wbar cubes
    const double *axr, *axc, *bx;
    double *gax;            // [no][nup][nup][kp]
    double *gbx;            // [no][no][nup][kp]
    const int *tuples;
    int t0;
    const int *ulist;       // gax: active x;  gbx: active y*no+z
    const int *uoff;        // CSR offsets into ent, per ulist entry
    const int *ent;         // (tl, sigma) pairs
    double *part;           // gbx: per (entry, split, chunk, gamma group) partial tiles
    int no, nup, nb, kp, nchunk, gt, ngg, nsplit;
    int ch0;                // gax: first K chunk of this launch (nchunk chunks from there)
};

// gAX[x](R0 16 x R1 16 rows, K chunk) += sum over entries (t,sigma) with t[xp] = x and over
// the gamma blocks G of Wbar_t(e_u = a, e_v = b, e_wf = g) BX[t[yp],t[zp]](g, K).
// Wave w owns rows b in {2w, 2w+1}, all 8 K tiles of the chunk.
// NCT: K tiles of 16 the work group computes (GCT for full chunks, fewer for the last chunk
// when kp is not a multiple of GCH, so no MFMA runs on K padding).
template <int NCT>
__global__ void __launch_bounds__(GTHR, 1) gax_kernel(GemmArgs a)
{
    extern __shared__ double lds[];
    double *WS = lds;                     // [2][WB_BLK]
    double *BS = lds + 2 * WB_BLK;        // [2][GCH * BS_STR]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int u = blockIdx.x;
    const int ch = a.ch0 + u % a.nchunk; u /= a.nchunk;
    const int R1 = u % a.nb; u /= a.nb;
    const int R0 = u % a.nb; u /= a.nb;
    const int ui = u;
    const int x = a.ulist[ui];
    const int e_beg = a.uoff[ui], e_end = a.uoff[ui + 1];
    const int K0 = ch * GCH;
    const long long cube = (long long)a.nb * a.nb * a.nb * 4096;
    const int nsteps = (e_end - e_beg) * a.nb;

    d4 acc[GBPW][NCT];
#pragma unroll
    for (int bb = 0; bb < GBPW; ++bb)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[bb][ct] = d4{0, 0, 0, 0};

    // staging registers: Wbar block 4096 = 512 x 8, BX slice 16 x 128 = 512 x 4
    double rw[8], rb[4];
    double rwgt = 1.0;                    // weight of the staged entry, applied when stashing
    auto fetch = [&](int step) {
        const int e = e_beg + step / a.nb, G = step % a.nb;
        const int tl = a.ent[2 * e], sg = a.ent[2 * e + 1];
        const int4 tp = reinterpret_cast<const int4 *>(a.tuples)[a.t0 + tl];
        const int tpos[3] = {tp.x, tp.y, tp.z};
        const Sigma S = SIG[sg & 7];
        rwgt = (sg & 8) ? 2.0 : 1.0;      // a merged pair of a degenerate tuple
        int bc[3];
        bc[S.u] = R0; bc[S.v] = R1; bc[S.wf] = G;
        const double *wsrc = a.wbar + tl * cube + (((long long)bc[0] * a.nb + bc[1]) * a.nb + bc[2]) * 4096;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double2 v = reinterpret_cast<const double2 *>(wsrc)[tid + k * GTHR];
            rw[2 * k] = v.x; rw[2 * k + 1] = v.y;
        }
        const double *bsrc = a.bx + (((long long)tpos[S.yp] * a.no + tpos[S.zp]) * a.nb + G) * (16LL * a.kp);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int idx = 2 * (tid + k * GTHR);           // element of the 128 x 16 slice
            const int Kl = idx >> 4;
            double2 v = {0.0, 0.0};
            if (K0 + Kl < a.kp) v = reinterpret_cast<const double2 *>(bsrc + (long long)(K0 + Kl) * 16)[(idx & 15) >> 1];
            rb[2 * k] = v.x; rb[2 * k + 1] = v.y;
        }
    };
    auto stash = [&](int buf) {
        double *W = WS + buf * WB_BLK;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = 2 * (tid + k * GTHR);           // (a,b,c) with c even
            const int aa = idx >> 8, b = (idx >> 4) & 15, c = idx & 15;
            W[aa * WB_STR_A + b * WB_STR_B + c] = rw[2 * k];
            W[aa * WB_STR_A + b * WB_STR_B + c + 1] = rw[2 * k + 1];
        }
        double *Bs = BS + buf * GCH * BS_STR;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int idx = 2 * (tid + k * GTHR);
            const int Kl = idx >> 4, gl = idx & 15;
            Bs[Kl * BS_STR + gl] = rb[2 * k] * rwgt;
            Bs[Kl * BS_STR + gl + 1] = rb[2 * k + 1] * rwgt;
        }
    };

    if (nsteps > 0) {
        fetch(0);
        stash(0);
    }
    __syncthreads();
    for (int step = 0; step < nsteps; ++step) {
        const int buf = step & 1;
        if (step + 1 < nsteps) fetch(step + 1);
        const int e = e_beg + step / a.nb;
        const Sigma S = SIG[a.ent[2 * e + 1] & 7];
        const int str[3] = {WB_STR_A, WB_STR_B, 1};
        const int su = str[S.u], sv = str[S.v], sw = str[S.wf];
        const double *W = WS + buf * WB_BLK;
        const double *Bs = BS + buf * GCH * BS_STR;
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) {
            const int g = 4 * kq + (lane >> 4);
            double af[GBPW];
#pragma unroll
            for (int bb = 0; bb < GBPW; ++bb) af[bb] = W[(lane & 15) * su + (wave * GBPW + bb) * sv + g * sw];
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) {
                const double bf = Bs[(ct * 16 + (lane & 15)) * BS_STR + g];
#pragma unroll
                for (int bb = 0; bb < GBPW; ++bb) acc[bb][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[bb], bf, acc[bb][ct], 0, 0, 0);
            }
        }
        if (step + 1 < nsteps) stash(buf ^ 1);
        __syncthreads();
    }
    // read-add-write of the output tile: rows a = row(lane,v), b = 2*wave+bb; column K
#pragma unroll
    for (int bb = 0; bb < GBPW; ++bb)
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
            const int K = K0 + ct * 16 + (lane & 15);
            if (K >= a.kp) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int al = Mfma<double>::row(lane, v);
                double *o = a.gax + (((long long)x * a.nup + R0 * TB + al) * a.nup + R1 * TB + wave * GBPW + bb) * a.kp + K;
                *o += acc[bb][ct][v];
            }
        }
}

// Partial gBX tile of one entry (t,sigma) of pair (y,z) = (t[yp],t[zp]): gamma tiles of group
// gg, K chunk ch, panels R0 in split rs (all R1), summed over the 256 (a,b) of a panel:
//   part(g, K) = sum Wbar_t(e_u = a, e_v = b, e_wf = g) AX[t[xp]](a,b,K).
// One work group per (entry, split, chunk, group), so pairs that collect many entries in a
// batch spread over many work groups; gbx_reduce_kernel adds the partials of a pair in a
// fixed order.  A panel is walked in 16 slices of fixed e_sa (sa = axis 0 when u or v is 0,
// else axis 1), each slice staged in LDS: Wbar (16 x gamma) and AX (16 x 128 K).  Wave w
// owns K tile w and all gamma tiles.
template <int GT>
__global__ void __launch_bounds__(GTHR, 1) gbx_kernel(GemmArgs a)
{
    extern __shared__ double lds[];
    constexpr int gstr = GT * 16 + 1;     // padded row of the staged Wbar slice
    double *WS = lds;                     // [2][16 * gstr]
    double *AS = lds + 2 * 16 * gstr;     // [2][GCH * BS_STR]
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int u = blockIdx.x;
    const int gg = u % a.ngg; u /= a.ngg;
    const int ch = u % a.nchunk; u /= a.nchunk;
    const int rs = u % a.nsplit; u /= a.nsplit;
    const int e = u;                      // entry (CSR position)
    const int r0_beg = rs * a.nb / a.nsplit, r0_end = (rs + 1) * a.nb / a.nsplit;
    const int K0 = ch * GCH;
    const int G0 = gg * GT;
    const int ngt = (G0 + GT <= a.nb) ? GT : a.nb - G0;   // tiles past ngt are zero
    const long long cube = (long long)a.nb * a.nb * a.nb * 4096;
    const int nsteps = (r0_end - r0_beg) * a.nb * 16;
    constexpr int nw = GT * 256;          // Wbar slice elements (zero past ngt tiles)

    d4 acc[GT];
#pragma unroll
    for (int i = 0; i < GT; ++i) acc[i] = d4{0, 0, 0, 0};

    // Staging registers.  The Wbar slice is read in memory order: with the slice axis sa
    // fixed, the block element offset is m>>4 along axis hi (1 if sa = 0, else 0) and m&15
    // along axis 2, so a wave reads 128-byte runs whatever sigma is.
    double2 rw[4];
    double ra[4];
    double rwgt = 1.0;                    // weight of the staged entry, applied when stashing
    int st_tr;                            // 1: the slice's (hi, 2) axes are (o, g); 0: (g, o)
    auto fetch = [&](int step) {
        int r = step;
        const int sidx = r & 15; r >>= 4;
        const int R1 = r % a.nb, R0 = r0_beg + r / a.nb;
        const int tl = a.ent[2 * e], sg = a.ent[2 * e + 1];
        const int4 tp = reinterpret_cast<const int4 *>(a.tuples)[a.t0 + tl];
        const int tpos[3] = {tp.x, tp.y, tp.z};
        const Sigma S = SIG[sg & 7];
        rwgt = (sg & 8) ? 2.0 : 1.0;      // a merged pair of a degenerate tuple
        const int sa = (S.u == 0 || S.v == 0) ? 0 : 1;
        const int hi = sa == 0 ? 1 : 0;
        const int oa = (sa == S.u) ? S.v : S.u;
        st_tr = (oa == hi) ? 1 : 0;
        int bc[3];
        bc[S.u] = R0; bc[S.v] = R1; bc[S.wf] = G0;
        const long long gstep = S.wf == 0 ? (long long)a.nb * a.nb : S.wf == 1 ? a.nb : 1;
        const double *wbase = a.wbar + tl * cube + (((long long)bc[0] * a.nb + bc[1]) * a.nb + bc[2]) * 4096 +
                              (sa == 0 ? sidx * 256 : sidx * 16);
        const int hstr = sa == 0 ? 16 : 256;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = tid + k * GTHR;               // double2 index in the slice
            double2 v = {0.0, 0.0};
            if (idx < nw / 2) {
                const int gt_l = idx >> 7, m = (idx & 127) * 2;
                if (gt_l < ngt)
                    v = *reinterpret_cast<const double2 *>(wbase + gt_l * gstep * 4096 + (m >> 4) * hstr + (m & 15));
            }
            rw[k] = v;
        }
        // AX slice: row layout (fixed alpha = e_u) or column layout (fixed beta = e_v); the
        // slab interleaves the outer indices 2w, 2w+1 per lane: element (K, o) of outer index
        // sidx sits at ((sidx/2 * kp/4 + K/4) * 64 + (K%4)*16 + o) * 2 + sidx%2 of the panel
        const double *src = (sa == S.u ? a.axr : a.axc) +
                            (((long long)tpos[S.xp] * a.nb + R0) * a.nb + R1) * (256LL * a.kp) +
                            (long long)(sidx >> 1) * (32LL * a.kp) + (sidx & 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = tid + k * GTHR;                 // (K, o) of the 128 x 16 slice
            const int Kl = e >> 4, o = e & 15, K = K0 + Kl;
            ra[k] = K < a.kp ? src[((long long)(K >> 2) * 64 + (K & 3) * 16 + o) * 2] : 0.0;
        }
    };
    auto stash = [&](int buf) {
        double *W = WS + buf * 16 * gstr;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = tid + k * GTHR;
            if (idx < nw / 2) {
                const int gt_l = idx >> 7, m = (idx & 127) * 2;
                const int h = m >> 4, l = m & 15;     // l, l+1 along axis 2
                if (st_tr) {                          // o = h, g = l, l+1
                    W[h * gstr + gt_l * 16 + l] = rw[k].x;
                    W[h * gstr + gt_l * 16 + l + 1] = rw[k].y;
                } else {                              // o = l, l+1, g = h
                    W[l * gstr + gt_l * 16 + h] = rw[k].x;
                    W[(l + 1) * gstr + gt_l * 16 + h] = rw[k].y;
                }
            }
        }
        double *As = AS + buf * GCH * BS_STR;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = tid + k * GTHR;
            As[(e >> 4) * BS_STR + (e & 15)] = ra[k] * rwgt;
        }
    };

    if (nsteps > 0) {
        fetch(0);
        stash(0);
    }
    __syncthreads();
    for (int step = 0; step < nsteps; ++step) {
        const int buf = step & 1;
        if (step + 1 < nsteps) fetch(step + 1);
        const double *W = WS + buf * 16 * gstr;
        const double *As = AS + buf * GCH * BS_STR;
#pragma unroll
        for (int kq = 0; kq < 4; ++kq) {
            const int o = 4 * kq + (lane >> 4);
            const double bf = As[(wave * 16 + (lane & 15)) * BS_STR + o];
#pragma unroll
            for (int i = 0; i < GT; ++i)
                acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(W[o * gstr + i * 16 + (lane & 15)], bf, acc[i], 0, 0, 0);
        }
        if (step + 1 < nsteps) stash(buf ^ 1);
        __syncthreads();
    }
    double *pt = a.part + (((long long)e * a.nsplit + rs) * a.nchunk + ch) * a.ngg * (GT * 16 * GCH) +
                 (long long)gg * (GT * 16 * GCH);
#pragma unroll
    for (int i = 0; i < GT; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            pt[(i * 16 + Mfma<double>::row(lane, v)) * GCH + wave * 16 + (lane & 15)] = acc[i][v];
}

// gbx[y][z][g][K] += sum over the entries of pair (y,z) (CSR order) and the panel splits of
// the partial tiles
__global__ void gbx_reduce_kernel(const double *__restrict__ part, double *__restrict__ gbx,
                                  const int *__restrict__ plist, const int *__restrict__ poff, int np, int nup,
                                  int kp, int nchunk, int ngg, int gtr, int nsplit)
{
    const long long per = (long long)nup * kp;
    const long long total = (long long)np * per;
    const long long tile = (long long)gtr * GCH;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int pi = (int)(idx / per);
        const int rem = (int)(idx % per);
        const int g = rem / kp, K = rem % kp;
        const int ch = K / GCH, kl = K % GCH, gg = g / gtr, gl = g % gtr;
        double s = 0.0;
        for (int e = poff[pi]; e < poff[pi + 1]; ++e)
            for (int rs = 0; rs < nsplit; ++rs)
                s += part[((((long long)e * nsplit + rs) * nchunk + ch) * ngg + gg) * tile + (long long)gl * GCH + kl];
        gbx[(long long)plist[pi] * per + rem] += s;
    }
}

// dvoe[q][r][X*16+i][Y*16+j] += sum over entries (tl, p) of pair (q,r) and blocks Z of
// slots[tl][p][X][Y][Z][i*16+j]
__global__ void dvoe_reduce_kernel(const double *__restrict__ slots, double *__restrict__ dvoe,
                                   const int *__restrict__ plist, const int *__restrict__ poff,
                                   const int *__restrict__ ent, int np, int no, int nup, int nb)
{
    const long long per = (long long)nup * nup;
    const long long total = (long long)np * per;
    const long long tstride = 6LL * nb * nb * nb * 256;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int pi = (int)(idx / per);
        const int rem = (int)(idx % per);
        const int xr = rem / nup, yc = rem % nup;
        const int X = xr >> 4, Y = yc >> 4, e = (xr & 15) * 16 + (yc & 15);
        double s = 0.0;
        for (int k = poff[pi]; k < poff[pi + 1]; ++k) {
            const double *sl = slots + ent[2 * k] * tstride + ((((long long)ent[2 * k + 1] * nb + X) * nb + Y) * nb) * 256 + e;
            for (int Z = 0; Z < nb; ++Z) s += sl[Z * 256];
        }
        dvoe[(long long)plist[pi] * per + rem] += s;
    }
}

// gAX / gBX / dVOE accumulators and v1 -> reference layouts
//   dVVVO(K,b,a,x) = gAX[x][a][b][K]                    (K < nu)
//   dT2(a,b,i,j)   = gAX[j][a][b][nu+i] + gBX[i][j][a][b]
//   dVM(l,g,z,y)   = -gBX[y][z][g][nu+l]
//   dVOE(x,y,q,r)  = dvoe[q][r][x][y]
//   dT1(o,a)       = 2 v1(a,o)
__global__ void unpack_grad_kernel(const double *__restrict__ gax, const double *__restrict__ gbx,
                                   const double *__restrict__ dvoep, const double *__restrict__ v1,
                                   double *__restrict__ dvvvo, double *__restrict__ dt2, double *__restrict__ dvm,
                                   double *__restrict__ dvoe, double *__restrict__ dt1, int no, int nu, int nup, int kp)
{
    const long long n_vvvo = (long long)nu * nu * nu * no;
    const long long n_t2 = (long long)nu * nu * no * no;
    const long long n_vm = (long long)no * nu * no * no;
    const long long n_t1 = (long long)no * nu;
    const long long total = n_vvvo + 2 * n_t2 + n_vm + n_t1;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        long long r = idx;
        if (r < n_vvvo) {
            if (dvvvo) {
                const int K = r % nu; long long q = r / nu;
                const int b = q % nu; q /= nu;
                const int a = q % nu; const int x = (int)(q / nu);
                dvvvo[r] = gax[(((long long)x * nup + a) * nup + b) * kp + K];
            }
            continue;
        }
        r -= n_vvvo;
        if (r < n_t2) {
            if (dt2) {
                const int a = r % nu; long long q = r / nu;
                const int b = q % nu; q /= nu;
                const int i = q % no; const int j = (int)(q / no);
                dt2[r] = gax[(((long long)j * nup + a) * nup + b) * kp + nu + i] +
                         gbx[(((long long)i * no + j) * nup + a) * kp + b];
            }
            continue;
        }
        r -= n_t2;
        if (r < n_vm) {
            if (dvm) {
                const int l = r % no; long long q = r / no;
                const int gm = q % nu; q /= nu;
                const int z = q % no; const int y = (int)(q / no);
                dvm[r] = -gbx[(((long long)y * no + z) * nup + gm) * kp + nu + l];
            }
            continue;
        }
        r -= n_vm;
        if (r < n_t2) {
            if (dvoe) {
                const int x = r % nu; long long q = r / nu;
                const int y = q % nu; q /= nu;
                const int qq = q % no; const int rr = (int)(q / no);
                dvoe[r] = dvoep[(((long long)qq * no + rr) * nup + x) * nup + y];
            }
            continue;
        }
        r -= n_t2;
        if (dt1) {
            const int o = r % no, a = (int)(r / no);
            dt1[r] = 2.0 * v1[a + (long long)nu * o];
        }
    }
}

template <int NCT>
static int launch_gax(const GemmArgs &a, long long grid, hipStream_t s)
{
    static bool attr_set = false;
    const size_t lds_bytes = sizeof(double) * (2 * WB_BLK + 2 * GCH * BS_STR);
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)gax_kernel<NCT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_bytes) != hipSuccess)
            return 2;
        attr_set = true;
    }
    gax_kernel<NCT><<<(unsigned)grid, GTHR, lds_bytes, s>>>(a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <int GT>
static int launch_gbx(const GemmArgs &a, long long grid, hipStream_t s)
{
    static bool attr_set = false;
    const size_t lds_bytes = sizeof(double) * (2 * 16 * (GT * 16 + 1) + 2 * GCH * BS_STR);
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)gbx_kernel<GT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_bytes) != hipSuccess)
            return 2;
        attr_set = true;
    }
    gbx_kernel<GT><<<(unsigned)grid, GTHR, lds_bytes, s>>>(a);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

extern "C" {

int ccx_hip_adj_lds_bytes(void) { return (int)(sizeof(double) * ADJ_LDS); }

size_t ccx_hip_adj_stage_doubles(void) { return (size_t)6 * BPW * 4 * NTHR; }

int ccx_hip_adj_items(const void *axr, const void *axc, const void *bx, const double *voep, const double *eh,
                      const double *ep, const double *t1, const int *tuples, const int *blocks, double *e_part,
                      double *v1_part, double *wbar, double *slots, double *wstage, unsigned int *counter, int no,
                      int nu, int nup, int kp, int ntup, int nblk, int t0, int tcount, int grid,
                      int form, const void *const *w32, int kps, double wscale, void *stream)
{
    // form 0: FP64 slabs axr/axc/bx; 1: FP32 slabs w32[] (K padded to kps, a multiple of 48);
    // 2: split BF16 slices w32[] (K padded to kps, a multiple of 64, W scaled by wscale)
    static bool attr_set[3] = {false, false, false};
    const size_t lds_bytes = sizeof(double) * ADJ_LDS;
    if (form < 0 || form > 2) return 3;
    if (form != 0 && w32 == nullptr) return 3;
    if ((form == 1 && kps % 48 != 0) || (form == 2 && kps % 64 != 0)) return 3;
    if (!attr_set[form]) {
        const void *fn = form == 0 ? (const void *)adj_item_kernel<0>
                       : form == 1 ? (const void *)adj_item_kernel<1> : (const void *)adj_item_kernel<2>;
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
            return 2;
        attr_set[form] = true;
    }
    const int sp = form == 2;
    AdjArgs ga;
    TriplesArgs &g = ga.f;
    g.axr = axr; g.axc = axc; g.bx = bx; g.voe = voep; g.eh = eh; g.ep = ep; g.tuples = tuples;
    g.blocks = blocks; g.e_part = e_part; g.v1_part = v1_part; g.wout = nullptr; g.counter = counter;
    g.no = no; g.nu = nu; g.nup = nup; g.nb = nup / TB; g.kp = kp; g.ntup = ntup; g.nblk = nblk;
    g.nitems = (long long)ntup * nblk;
    g.block_major = 0;
    if (form == 1) { g.axr = w32[0]; g.axc = w32[1]; g.bx = w32[2]; }
    g.axr_s = sp ? w32[0] : nullptr;
    g.axc_s = sp ? w32[1] : nullptr;
    g.bx_s = sp ? w32[2] : nullptr;
    g.axr_e = g.axc_e = g.bx_e = nullptr;
    g.kps = kps;
    g.wscale = sp ? wscale : 1.0;
    ga.t1 = t1; ga.wbar = wbar; ga.slots = slots; ga.wstage = wstage; ga.t0 = t0; ga.tcount = tcount;
    const long long nlocal = (long long)tcount * nblk;
    if (nlocal == 0) return 0;
    if (grid <= 0) grid = 256;
    if (grid > nlocal) grid = (int)nlocal;
    hipStream_t s = (hipStream_t)stream;
    if (hipMemsetAsync(counter, 0, sizeof(unsigned int), s) != hipSuccess) return 1;
    if (form == 2) adj_item_kernel<2><<<grid, NTHR, lds_bytes, s>>>(ga);
    else if (form == 1) adj_item_kernel<1><<<grid, NTHR, lds_bytes, s>>>(ga);
    else adj_item_kernel<0><<<grid, NTHR, lds_bytes, s>>>(ga);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_gax(const double *wbar, const double *bx, double *gax, const int *tuples, int t0, const int *xlist,
                const int *xoff, const int *ent, int nx, int no, int nup, int kp, void *stream)
{
    if (nx == 0) return 0;
    GemmArgs a = {};
    a.wbar = wbar; a.bx = bx; a.gax = gax; a.tuples = tuples; a.t0 = t0;
    a.ulist = xlist; a.uoff = xoff; a.ent = ent;
    a.no = no; a.nup = nup; a.nb = nup / TB; a.kp = kp;
    hipStream_t s = (hipStream_t)stream;
    // full K chunks, then the last partial chunk with only its K tiles
    const int nfull = kp / GCH, tail = (kp % GCH + 15) / 16;
    int rc = 0;
    if (nfull > 0) {
        a.ch0 = 0;
        a.nchunk = nfull;
        rc = launch_gax<GCT>(a, (long long)nx * a.nb * a.nb * nfull, s);
    }
    if (rc == 0 && tail > 0) {
        a.ch0 = nfull;
        a.nchunk = 1;
        const long long grid = (long long)nx * a.nb * a.nb;
        switch (tail) {
        case 1: rc = launch_gax<1>(a, grid, s); break;
        case 2: rc = launch_gax<2>(a, grid, s); break;
        case 3: rc = launch_gax<3>(a, grid, s); break;
        case 4: rc = launch_gax<4>(a, grid, s); break;
        case 5: rc = launch_gax<5>(a, grid, s); break;
        case 6: rc = launch_gax<6>(a, grid, s); break;
        case 7: rc = launch_gax<7>(a, grid, s); break;
        default: rc = launch_gax<GCT>(a, grid, s); break;   // kp % GCH > 112
        }
    }
    return rc;
}

void ccx_hip_gbx_geometry(int nup, int kp, int *nchunk, int *ngg, int *gtr)
{
    const int nb = nup / TB;
    *ngg = (nb + 13) / 14;                // gamma groups of at most 14 tiles
    *gtr = 16 * ((nb + *ngg - 1) / *ngg); // rows of a partial tile
    *nchunk = (kp + GCH - 1) / GCH;
}

size_t ccx_hip_gbx_part_doubles(int nup, int kp, int nent, int nsplit)
{
    int nchunk, ngg, gtr;
    ccx_hip_gbx_geometry(nup, kp, &nchunk, &ngg, &gtr);
    return (size_t)nent * nsplit * nchunk * ngg * gtr * GCH;
}

int ccx_hip_gbx(const double *wbar, const double *axr, const double *axc, double *gbx, double *part,
                const int *tuples, int t0, const int *plist, const int *poff, const int *ent, int npair, int nent,
                int nsplit, int no, int nup, int kp, void *stream)
{
    const int nb = nup / TB;
    int nchunk, ngg, gtr;
    ccx_hip_gbx_geometry(nup, kp, &nchunk, &ngg, &gtr);
    const int gt = gtr / 16;
    if (npair == 0 || nent == 0) return 0;
    GemmArgs a = {};
    a.wbar = wbar; a.axr = axr; a.axc = axc; a.gbx = gbx; a.part = part; a.tuples = tuples; a.t0 = t0;
    a.ulist = plist; a.uoff = poff; a.ent = ent;
    a.no = no; a.nup = nup; a.nb = nb; a.kp = kp; a.nchunk = nchunk; a.gt = gt; a.ngg = ngg; a.nsplit = nsplit;
    const long long grid = (long long)nent * nsplit * nchunk * ngg;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    switch (gt) {
    case 1: rc = launch_gbx<1>(a, grid, s); break;
    case 2: rc = launch_gbx<2>(a, grid, s); break;
    case 3: rc = launch_gbx<3>(a, grid, s); break;
    case 4: rc = launch_gbx<4>(a, grid, s); break;
    case 5: rc = launch_gbx<5>(a, grid, s); break;
    case 6: rc = launch_gbx<6>(a, grid, s); break;
    case 7: rc = launch_gbx<7>(a, grid, s); break;
    case 8: rc = launch_gbx<8>(a, grid, s); break;
    case 9: rc = launch_gbx<9>(a, grid, s); break;
    case 10: rc = launch_gbx<10>(a, grid, s); break;
    case 11: rc = launch_gbx<11>(a, grid, s); break;
    case 12: rc = launch_gbx<12>(a, grid, s); break;
    case 13: rc = launch_gbx<13>(a, grid, s); break;
    default: rc = launch_gbx<14>(a, grid, s); break;
    }
    if (rc != 0) return rc;
    const long long total = (long long)npair * nup * kp;
    long long g2 = (total + 255) / 256;
    if (g2 > 65536) g2 = 65536;
    gbx_reduce_kernel<<<(unsigned)g2, 256, 0, s>>>(part, gbx, plist, poff, npair, nup, kp, nchunk, ngg, gtr, nsplit);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_dvoe_reduce(const double *slots, double *dvoep, const int *plist, const int *poff, const int *ent,
                        int np, int no, int nup, void *stream)
{
    if (np == 0) return 0;
    const long long total = (long long)np * nup * nup;
    long long grid = (total + 255) / 256;
    if (grid > 65536) grid = 65536;
    dvoe_reduce_kernel<<<(unsigned)grid, 256, 0, (hipStream_t)stream>>>(slots, dvoep, plist, poff, ent, np, no, nup,
                                                                       nup / TB);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_unpack_grad(const double *gax, const double *gbx, const double *dvoep, const double *v1, double *dvvvo,
                        double *dt2, double *dvm, double *dvoe, double *dt1, int no, int nu, int nup, int kp,
                        void *stream)
{
    const long long total = (long long)nu * nu * nu * no + 2LL * nu * nu * no * no + (long long)no * nu * no * no +
                            (long long)no * nu;
    long long grid = (total + 255) / 256;
    if (grid > 65536) grid = 65536;
    unpack_grad_kernel<<<(unsigned)grid, 256, 0, (hipStream_t)stream>>>(gax, gbx, dvoep, v1, dvvvo, dt2, dvm, dvoe, dt1,
                                                                       no, nu, nup, kp);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

int ccx_hip_zero(void *p, size_t bytes, void *stream)
{
    if (bytes == 0) return 0;
    return hipMemsetAsync(p, 0, bytes, (hipStream_t)stream) == hipSuccess ? 0 : 1;
}

}  // extern "C"
