// gf_btdsystem_noalloc: the block-tridiagonal solve of gf_btdsystem for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  Inputs and outputs are the caller's device buffers; the
// factor workspace (the copies of A, B, C that cyclic reduction reduces in
// place, B^-1, X, Y, the saved couplings, the right-hand-side work vector, the
// level table and the failure flag) lives in `scratch`, a file-scope device
// array of fixed size that every GPU the library runs on holds until the
// process ends.  Nothing here calls hipMalloc, hipFree or an allocating
// library routine, and nothing waits on the host: the work is enqueued on the
// caller's stream, fills are kernels (no memset nodes in a captured graph).
//
// Every call on one GPU shares that GPU's scratch and leaves its contents
// behind; whatever must start at zero (X, Y, the flag, the couplings past
// either end) is cleared by each call.  Calls that overlap on one GPU (other
// streams, other threads) race on the scratch.
#include "gf_internal.h"

#include <mutex>

namespace {

// Capacity: nblk * m * m <= 2^17 block entries (so n = nblk m <= 2^17 rows).
constexpr size_t NA_ELEMS = (size_t)1 << 17;
// A|B|C, Bi, X|Y (6 E), Lsv and Rsv (at most 2 E), the work vector (at most E),
// the level table and the flag, each rounded up to 32 doubles: 9.4 MB.
constexpr size_t NA_SCRATCH = 9 * NA_ELEMS + 512;

static __device__ __align__(256) double scratch[NA_SCRATCH];

constexpr int NA_MAX_DEV = 64;
std::mutex na_mu;
double *na_ptr[NA_MAX_DEV] = {};

// Device address of this GPU's scratch (hipGetSymbolAddress once per device,
// with that device current); the array's name is a host-side stand-in only.
int scratch_of(int dev, double **out)
{
    if (dev < 0 || dev >= NA_MAX_DEV) {
        gf_set_error("gf_btdsystem_noalloc: device ordinal out of range");
        return GF_E_ARG;
    }
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_ptr[dev]) {
        void *p = nullptr;
        GF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        na_ptr[dev] = (double *)p;
    }
    *out = na_ptr[dev];
    return GF_OK;
}

struct LevelTable {
    long long v[32];
    int count;
};

__global__ void level_table_kernel(LevelTable t, long long *dst)
{
    const int i = threadIdx.x;
    if (i < t.count) dst[i] = t.v[i];
}

__global__ void copy_flag_kernel(const int *src, int *dst)
{
    if (threadIdx.x == 0) *dst = (*src & GF_FLAG_FAIL) ? 1 : 0;
}

}  // namespace

extern "C" int gf_btdsystem_noalloc(int m, int nblk, const double *A, const double *B, const double *C,
                                    const double *b, double *x, int *info_fail, void *stream)
{
    if (m < 1 || m > 32 || nblk < 1 || !A || !B || !C || !b || !x || !info_fail) return GF_E_ARG;
    const size_t bs = (size_t)m * m, E = (size_t)nblk * bs, n = (size_t)nblk * m;
    if (E > NA_ELEMS) {
        gf_set_error("gf_btdsystem_noalloc: nblk * m * m exceeds the scratch capacity of 131072");
        return GF_E_CAPACITY;
    }
    int dev = 0;
    GF_HIP(hipGetDevice(&dev));
    double *sp = nullptr;
    GF_TRY(scratch_of(dev, &sp));
    hipStream_t s = (hipStream_t)stream;
    gf_btd t;
    GF_TRY(gf_btd_plan(t, m, nblk, 8));
    if ((size_t)t.n_save > (size_t)nblk || t.lvl_save.size() + 1 > 32) {
        gf_set_error("gf_btdsystem_noalloc: level table exceeds the scratch layout");
        return GF_E_CAPACITY;
    }
    // carve the scratch: A|B|C contiguous (the kernels take the block kind as an
    // offset from A), then Bi, X, Y, Lsv, Rsv, the work vector, table and flag,
    // each of the separate arrays starting on a 256-byte boundary
    size_t off = 0;
    auto take = [&](size_t count) {
        double *p = sp + off;
        off += (count + 31) / 32 * 32;
        return p;
    };
    t.A = take(3 * E);
    t.B = (double *)t.A + E;
    t.C = (double *)t.A + 2 * E;
    t.Bi = take(E);
    t.X = take(2 * E);  // X and Y adjacent: one fill clears both
    t.Y = (double *)t.X + E;
    const size_t sv = std::max<size_t>(1, (size_t)t.n_save) * bs;
    t.Lsv = take(sv);
    t.Rsv = take(sv);
    double *work = take(n);
    t.d_save = (long long *)take(32);
    int *flag = (int *)take(1);
    if (off > NA_SCRATCH) {
        gf_set_error("gf_btdsystem_noalloc: scratch layout overflow");
        return GF_E_CAPACITY;
    }
    // inputs into the scratch, couplings past either end cleared
    GF_HIP(hipMemcpyAsync(t.A, A, E * 8, hipMemcpyDeviceToDevice, s));
    GF_HIP(hipMemcpyAsync(t.B, B, E * 8, hipMemcpyDeviceToDevice, s));
    GF_HIP(hipMemcpyAsync(t.C, C, E * 8, hipMemcpyDeviceToDevice, s));
    GF_TRY(gf_launch_fill(s, (double *)t.A, bs, 0.0));
    GF_TRY(gf_launch_fill(s, (double *)t.C + (E - bs), bs, 0.0));
    // X and Y of rows never eliminated with a coupling must read as zero
    GF_TRY(gf_launch_fill(s, (double *)t.X, 2 * E, 0.0));
    GF_TRY(gf_launch_fill_bytes(s, flag, 4));
    LevelTable lt{};
    lt.count = (int)t.lvl_save.size() + 1;
    for (size_t i = 0; i < t.lvl_save.size(); i++) lt.v[i] = t.lvl_save[i];
    lt.v[t.lvl_save.size()] = t.n_save;
    hipLaunchKernelGGL(level_table_kernel, dim3(1), dim3(64), 0, s, lt, t.d_save);
    GF_HIP(hipGetLastError());
    const int l = gf_btd_factor_solve(t, s, flag, b, x, (int)n, work);
    if (l < 0) return l;
    hipLaunchKernelGGL(copy_flag_kernel, dim3(1), dim3(64), 0, s, (const int *)flag, info_fail);
    GF_HIP(hipGetLastError());
    return GF_OK;
}
