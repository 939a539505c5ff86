// st_inject.hip -- orders parcels injected into a sorted cloud by cell.
//
// Injected parcels are stepped as the cloud's unsorted tail, one thread per parcel,
// until a re-sort takes them into the bricks.  Each tail deposit is an FP64 atomic on
// a cell of the source; parcels injected together crowd into the cells next to the
// inlet, and in injection order a wave's 64 lanes touch as many different cells.
// Ordered by cell, neighbouring lanes share cells, and the tail's waves sum their
// deposits per run of equal cells before the atomics (wave_deposit in st_step.hip).
// So each batch of injected parcels is sorted by linear cell index (the cloud's sorted
// mesh, x fastest) on the device: keys with indices, a radix sort of the pairs, a
// gather into the second buffer set and a copy back.
#include "st_handles.h"

#include <hipcub/device/device_radix_sort.hpp>

namespace {

template <class T>
__global__ void k_cell_keys(st::Soa<T> s, uint32_t k, st::Geo g, uint32_t *keys, uint32_t *idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const T p[3] = {s.x[i], s.y[i], s.z[i]};
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) {
        const double v = floor(((double)p[a] - g.o[a]) * g.rd[a]);
        c[a] = v < 0.0 ? 0u : (v >= (double)g.n[a] ? (uint32_t)(g.n[a] - 1) : (uint32_t)v);
    }
    keys[i] = c[2] * g.nxny + c[1] * (uint32_t)g.n[0] + c[0];
    idx[i] = i;
}

// out[i] = in[perm[i]] for the parcel arrays and their upload-order indices.
template <class T>
__global__ void k_gather_range(st::Soa<T> in, const int32_t *pid_in, const uint32_t *perm, st::Soa<T> out,
                               int32_t *pid_out, uint32_t k) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const uint32_t j = perm[i];
    out.d[i] = in.d[j];
    out.x[i] = in.x[j];
    out.y[i] = in.y[j];
    out.z[i] = in.z[j];
    out.u[i] = in.u[j];
    out.v[i] = in.v[j];
    out.w[i] = in.w[j];
    if (in.rho) const_cast<T *>(out.rho)[i] = in.rho[j];
    pid_out[i] = pid_in[j];
}

// Injected parcels from pinned host memory (`stage`: nf arrays of k values, d x y z u v w
// [rho]) into the cloud's arrays at [n0, n0 + k), each with its upload-order index: one
// kernel reading over PCIe, instead of a copy per array on the DMA engine, each of which
// hands over to and from the compute queue.
template <class T>
__global__ void k_inject_from_host(const T *__restrict__ stage, uint32_t k, st::Soa<T> out, int32_t *pid,
                                   int32_t first) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    out.d[i] = stage[i];
    out.x[i] = stage[(size_t)k + i];
    out.y[i] = stage[2 * (size_t)k + i];
    out.z[i] = stage[3 * (size_t)k + i];
    out.u[i] = stage[4 * (size_t)k + i];
    out.v[i] = stage[5 * (size_t)k + i];
    out.w[i] = stage[6 * (size_t)k + i];
    if (out.rho) const_cast<T *>(out.rho)[i] = stage[7 * (size_t)k + i];
    pid[i] = first + (int32_t)i;
}

template <class T>
st::Soa<T> offset(st::Soa<T> s, int64_t n0) {
    s.d += n0;
    s.x += n0;
    s.y += n0;
    s.z += n0;
    s.u += n0;
    s.v += n0;
    s.w += n0;
    if (s.rho) s.rho += n0;
    return s;
}

template <class T>
hipError_t sort_range(st_cloud *c, int64_t n0, uint32_t k, const st::Geo &g, void *tmp, size_t sort_bytes,
                      uint32_t *keys, uint32_t *idx, int bits, hipStream_t s) {
    using namespace st::host;
    const st::Soa<T> cur = offset(soa<T>(c, c->cur), n0), other = offset(soa<T>(c, 1 - c->cur), n0);
    int32_t *pid_cur = c->pid[c->cur] + n0, *pid_other = c->pid[1 - c->cur] + n0;
    const unsigned grid = (k + 255) / 256;
    k_cell_keys<T><<<grid, 256, 0, s>>>(cur, k, g, keys, idx);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, keys + k, idx, idx + k, (int)k, 0, bits, s);
    if (e == hipSuccess) {
        k_gather_range<T><<<grid, 256, 0, s>>>(cur, pid_cur, idx + k, other, pid_other, k);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = st::launch_copy_back<T>(other, pid_other, cur, pid_cur, k, nullptr, nullptr, s);
    return e;
}

}  // namespace

namespace st {
namespace host {

int copy_injected(st_cloud *c, const void *stage, int64_t n0, int64_t k) {
    const unsigned grid = (unsigned)((k + 255) / 256);
    hipStream_t s = c->ctx->stream;
    if (c->prec == ST_FP64)
        k_inject_from_host<double><<<grid, 256, 0, s>>>((const double *)stage, (uint32_t)k,
                                                        offset(soa<double>(c, c->cur), n0), c->pid[c->cur] + n0,
                                                        (int32_t)n0);
    else
        k_inject_from_host<float><<<grid, 256, 0, s>>>((const float *)stage, (uint32_t)k,
                                                       offset(soa<float>(c, c->cur), n0), c->pid[c->cur] + n0,
                                                       (int32_t)n0);
    ST_HIP(hipGetLastError());
    return ST_OK;
}

int sort_injected(st_cloud *c, int64_t n0, int64_t k) {
    if (k < 2 || !c->sorted) return ST_OK;
    st_mesh m;
    for (int a = 0; a < 3; ++a) {
        m.n[a] = c->sort_n[a];
        m.origin[a] = c->sort_o[a];
        m.ending[a] = c->sort_e[a];
    }
    const Geo g = make_geo(m);
    const uint64_t ncell = (uint64_t)g.nxny * (uint64_t)g.n[2];
    int bits = 1;
    while (bits < 32 && (1ull << bits) < ncell) ++bits;
    ST_TRY(alloc_set(c, 1 - c->cur));
    const uint32_t kk = (uint32_t)k;
    size_t sort_bytes = 0;
    ST_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr, (int)kk, 0, bits,
                                              c->ctx->stream));
    // keys and indices, in and out, then the radix sort's scratch: kept in the cloud and
    // grown on demand (once the stream is done with the old buffer)
    const size_t pairs = (size_t)4 * kk * sizeof(uint32_t);
    const size_t need = pairs + sort_bytes;
    if (need > c->inj_cap) {
        ST_HIP(hipStreamSynchronize(c->ctx->stream));
        if (c->inj_buf) ST_HIP(hipFree(c->inj_buf));
        c->inj_buf = nullptr;
        c->inj_cap = 0;
        if (hipMalloc(&c->inj_buf, need) != hipSuccess)
            return fail(ST_E_NOMEM, "cannot order %lld injected parcels by cell", (long long)k);
        c->inj_cap = need;
    }
    uint32_t *keys = (uint32_t *)c->inj_buf, *idx = keys + 2 * (size_t)kk;
    void *tmp = (char *)c->inj_buf + pairs;
    hipStream_t s = c->ctx->stream;
    const hipError_t e = c->prec == ST_FP64 ? sort_range<double>(c, n0, kk, g, tmp, sort_bytes, keys, idx, bits, s)
                                            : sort_range<float>(c, n0, kk, g, tmp, sort_bytes, keys, idx, bits, s);
    if (e != hipSuccess) return fail(ST_E_DEVICE, "ordering injected parcels by cell: %s", hipGetErrorString(e));
    return ST_OK;
}

}  // namespace host
}  // namespace st
