// ccx_energy_static: the FP64 (T) solve of ccx_load_device + ccx_energy with every work buffer
// (operand slabs, padded VOE and eps_virtual, tuple and block tables, per-item partials, queue
// counter, NaN flag) carved out of one file-scope device array instead of hipMalloc'd memory.
// For callers that cannot allocate device memory.  The kernels are the ones the context path
// launches, with the same grid, so the results are bitwise those of the context path.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdint.h>

#include "../../include/libcc_mi355x.h"
extern "C" {
#include "ccx_dev.h"
}

// One copy per GPU, held until the process ends.  Host code never passes this name to a kernel
// or a copy: on the host it is only a stand-in for the symbol; the device address of each GPU's
// copy comes from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[CCX_STATIC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
constexpr int TBLK = 16;   // virtual block edge of the persistent kernel

std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES];   // device address of scratch, per device ordinal

// Device address of the current device's scratch array; looked up once per device, with that
// device current.
int scratch_address(void **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return CCX_ERR_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return CCX_ERR_DEVICE;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CCX_OK;
}

// Tuple table (i>j>k with i outermost, then (i,j,j), (i,i,j) per i>j; kind 1/2/3) and block-orbit
// table (A<=B<=C, multiplicity 1/2/6), the orders ccx_create builds on the host.  One thread:
// the tables of a problem that fits the scratch array have at most a few ten thousand entries.
__global__ void static_tables_kernel(int no, int nb, int *__restrict__ tuples, int *__restrict__ blocks)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int n = 0;
    for (int i = 0; i < no; ++i)
        for (int j = 0; j < i; ++j)
            for (int k = 0; k < j; ++k) {
                tuples[4 * n] = i; tuples[4 * n + 1] = j; tuples[4 * n + 2] = k; tuples[4 * n + 3] = 1;
                ++n;
            }
    for (int i = 0; i < no; ++i)
        for (int j = 0; j < i; ++j) {
            tuples[4 * n] = i; tuples[4 * n + 1] = j; tuples[4 * n + 2] = j; tuples[4 * n + 3] = 2;
            tuples[4 * n + 4] = i; tuples[4 * n + 5] = i; tuples[4 * n + 6] = j; tuples[4 * n + 7] = 3;
            n += 2;
        }
    int m = 0;
    for (int a = 0; a < nb; ++a)
        for (int b = a; b < nb; ++b)
            for (int c = b; c < nb; ++c) {
                blocks[4 * m] = a; blocks[4 * m + 1] = b; blocks[4 * m + 2] = c;
                blocks[4 * m + 3] = (a == b && b == c) ? 6 : (a == b || b == c) ? 2 : 1;
                ++m;
            }
}

// eps_virtual padded to nup with finite values (padded virtual indices carry W = 0)
__global__ void static_pad_ep_kernel(const double *__restrict__ ep, double *__restrict__ epp, int nu, int nup)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < nup) epp[a] = a < nu ? ep[a] : 1.0e3;
}

__global__ void static_fill_zero_kernel(double *__restrict__ p, long long n)
{
    for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        p[q] = 0.0;
}

// Bump allocator over the scratch array: 256-byte aligned pieces.
struct Carve {
    size_t off = 0;
    unsigned char *base = nullptr;
    template <class T> T *take(long long count)
    {
        T *p = reinterpret_cast<T *>(base + off);
        off += ((size_t)(count > 0 ? count : 1) * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

struct Pieces {
    double *axr, *axc, *bx, *voep, *epp, *epart, *v1part, *v1t;
    int *tuples, *blocks;
    unsigned int *counter, *flag;
};

Pieces carve(Carve &cv, long long no, long long nu, long long nup, long long nb, long long kp, long long ntup,
             long long nblk)
{
    Pieces p;
    const long long nitems = ntup * nblk;
    p.axr = cv.take<double>(no * nb * nb * kp * 256);
    p.axc = cv.take<double>(no * nb * nb * kp * 256);
    p.bx = cv.take<double>(no * no * nb * kp * 16);
    p.voep = cv.take<double>(no * no * nup * nup);
    p.epp = cv.take<double>(nup);
    p.epart = cv.take<double>(nitems);
    p.v1part = cv.take<double>(nitems * 144);
    p.v1t = cv.take<double>(ntup * nu * 3);
    p.tuples = cv.take<int>(4 * ntup);
    p.blocks = cv.take<int>(4 * nblk);
    p.counter = cv.take<unsigned int>(1);
    p.flag = cv.take<unsigned int>(1);
    return p;
}

}  // namespace

extern "C" int ccx_energy_static(int64_t no, int64_t nu, const double *eh, const double *ep, const double *t1,
                                 const double *t2, const double *vm, const double *voe, const double *vvvo,
                                 double *e, double *v1, void *stream)
{
    if (!eh || !ep || !t1 || !t2 || !vm || !voe || !vvvo || !e || !v1) return CCX_ERR_ARG;
    if (no < 1 || nu < 1 || no > 2048 || nu > 4096) return CCX_ERR_ARG;
    const long long nup = (nu + TBLK - 1) / TBLK * TBLK, nb = nup / TBLK, kp = (nu + no + 3) / 4 * 4;
    const long long ntup = no * (no - 1) * (no - 2) / 6 + no * (no - 1);
    const long long nblk = nb * (nb + 1) * (nb + 2) / 6;
    {   // size check on the host, before any device work
        Carve probe;
        carve(probe, no, nu, nup, nb, kp, ntup, nblk);
        if (probe.off > (size_t)CCX_STATIC_SCRATCH_BYTES) return CCX_ERR_SCRATCH;
    }
    void *base = nullptr;
    int rc = scratch_address(&base);
    if (rc != CCX_OK) return rc;
    Carve cv;
    cv.base = static_cast<unsigned char *>(base);
    const Pieces p = carve(cv, no, nu, nup, nb, kp, ntup, nblk);
    hipStream_t s = (hipStream_t)stream;

    // the contract of ccx_load_device: NaN/Inf anywhere is CCX_ERR_INPUT
    const double *arr[7] = {eh, ep, t1, t2, vm, voe, vvvo};
    const size_t n[7] = {(size_t)no, (size_t)nu, (size_t)(no * nu), (size_t)(nu * nu * no * no),
                         (size_t)(no * nu * no * no), (size_t)(nu * nu * no * no), (size_t)(nu * nu * nu * no)};
    int bad = 0;
    if (ccx_hip_nonfinite(arr, n, 7, p.flag, &bad, stream) != 0) return CCX_ERR_DEVICE;
    if (bad) return CCX_ERR_INPUT;

    if (ntup == 0) {   // no = 1: no tuple, E(T) = 0 and v1 = 0
        static_fill_zero_kernel<<<1, 256, 0, s>>>(e, 2);
        static_fill_zero_kernel<<<(int)((nu * no + 255) / 256), 256, 0, s>>>(v1, nu * no);
        return hipGetLastError() == hipSuccess ? CCX_OK : CCX_ERR_DEVICE;
    }
    static_tables_kernel<<<1, 1, 0, s>>>((int)no, (int)nb, p.tuples, p.blocks);
    static_pad_ep_kernel<<<(int)((nup + 255) / 256), 256, 0, s>>>(ep, p.epp, (int)nu, (int)nup);
    if (hipGetLastError() != hipSuccess) return CCX_ERR_DEVICE;
    if (ccx_hip_pack(0, vvvo, t2, vm, voe, p.axr, p.axc, p.bx, p.voep, (int)no, (int)nu, (int)nup, (int)kp,
                     stream) != 0)
        return CCX_ERR_DEVICE;
    int cus = 256;
    if (ccx_hip_cu_count(&cus) != 0 || cus < 1) cus = 256;
    const long long nitems = ntup * nblk;
    const int grid = (int)(nitems < cus ? nitems : cus);
    if (ccx_hip_triples(0, p.axr, p.axc, p.bx, p.voep, eh, p.epp, p.tuples, p.blocks, p.epart, p.v1part, nullptr,
                        p.counter, (int)no, (int)nu, (int)nup, (int)kp, (int)ntup, (int)nblk, grid, nullptr, 0, 1.0,
                        stream) != 0)
        return CCX_ERR_DEVICE;
    if (ccx_hip_finalize(p.epart, p.v1part, p.tuples, p.blocks, t1, p.v1t, v1, e, (int)no, (int)nu, (int)ntup,
                         (int)nblk, stream) != 0)
        return CCX_ERR_DEVICE;
    return CCX_OK;
}
