// asan_driver.cpp -- drives every entry point of the C-ABI (include/scaletrack.h) from a
// plain C++ host program, for a build of the library's host code under AddressSanitizer
// (tests/native/Makefile).  The device kernels are the library's own; only the host side
// is instrumented, so ASan reports any out-of-bounds or use-after-free access of the
// host code: handle lifetimes, staging buffers, the copies of parcel and cell arrays,
// the batch and adjoint tables, and the error paths.
//
// It runs small clouds through: both precision forms and methods; the wall, periodic and
// channel forms; injection, compaction, reserve and re-sorts; a captured graph; the
// no-allocation step; a batch with its graph and the autotuner; the adjoint; a list of
// refused calls; and it destroys everything, including in an order that leaves a
// destroyed context alive behind its clouds.  Exit status 0 and "asan_driver: ok" on
// success; any failed call prints its message and exits 1.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/scaletrack.h"

namespace {

int g_fail = 0;

void expect(int got, int want, const char *what) {
    if (got != want) {
        std::fprintf(stderr, "asan_driver: %s returned %d (wanted %d): %s\n", what, got, want, st_last_error());
        g_fail = 1;
        std::exit(1);
    }
}
#define OK(call) expect((call), ST_OK, #call)

// SplitMix64 stream of uniform doubles in [0, 1).
struct Rng {
    uint64_t s;
    double next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

template <class T>
struct Host {
    std::vector<T> d, x, y, z, u, v, w;
    void fill(int64_t n, const st_mesh &m, uint64_t seed) {
        Rng r{seed};
        d.resize(n), x.resize(n), y.resize(n), z.resize(n), u.assign(n, 0), v.assign(n, 0), w.assign(n, 0);
        for (int64_t i = 0; i < n; ++i) {
            d[i] = (T)(30e-6 * (0.5 + r.next()));
            T *p[3] = {&x[i], &y[i], &z[i]};
            for (int a = 0; a < 3; ++a) {
                const double L = m.ending[a] - m.origin[a];
                T c = (T)(m.origin[a] + L * (0.02 + 0.96 * r.next()));
                *p[a] = c;
            }
        }
    }
};

std::vector<double> taylor_green(const st_mesh &m) {
    const int64_t nc = (int64_t)m.n[0] * m.n[1] * m.n[2];
    std::vector<double> U(nc * 3);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < m.n[2]; ++k)
        for (int j = 0; j < m.n[1]; ++j)
            for (int i = 0; i < m.n[0]; ++i) {
                const double X = (i + 0.5) / m.n[0], Y = (j + 0.5) / m.n[1], Z = (k + 0.5) / m.n[2];
                const int64_t c = ((int64_t)k * m.n[1] + j) * m.n[0] + i;
                U[3 * c] = std::sin(2 * pi * X) * std::cos(2 * pi * Y) * std::cos(2 * pi * Z);
                U[3 * c + 1] = -std::cos(2 * pi * X) * std::sin(2 * pi * Y) * std::cos(2 * pi * Z);
                U[3 * c + 2] = 0.2 * std::sin(2 * pi * Z);
            }
    return U;
}

template <class T>
void forms(st_ctx *ctx, int prec) {
    const st_mesh m{{20, 12, 9}, {0.0, 0.0, 0.0}, {1.0, 0.6, 0.45}};
    const int64_t n = 5000, ncell = 20 * 12 * 9;
    std::vector<double> U = taylor_green(m);
    std::vector<double> src(ncell * 3);
    for (int bnd : {ST_WALL, ST_PERIODIC, ST_CHANNEL})
        for (int method : {ST_METHOD_PLAIN, ST_METHOD_SORTED}) {
            st_field *f = nullptr;
            st_cloud *c = nullptr;
            OK(st_field_create(ctx, &m, prec, &f));
            OK(st_field_upload_U(f, U.data(), 8));
            OK(st_field_zero_source(f));
            OK(st_cloud_create(ctx, prec, n, 1e-3, 1000.0, 1.0, 0.0, 0, &c));
            Host<T> h;
            h.fill(n, m, 7 + bnd * 2 + method);
            OK(st_cloud_upload(c, h.d.data(), h.x.data(), h.y.data(), h.z.data(), h.u.data(), h.v.data(),
                               h.w.data(), nullptr));
            st_params p;
            st_params_default(&p);
            p.boundary = bnd;
            p.method = method;
            p.brick[0] = 8, p.brick[1] = 4, p.brick[2] = 2;
            p.substeps_per_pass = 1;
            p.resort_fraction = 0.0;  // the device re-sorts behind every step
            double bbox[6];
            for (int step = 0; step < 3; ++step) OK(st_evolve(ctx, c, f, 1e-3, 10, &p, bbox));
            p.substeps_per_pass = 0;
            OK(st_evolve(ctx, c, f, 1e-3, 10, &p, nullptr));
            if (method == ST_METHOD_SORTED) {
                OK(st_evolve_noalloc(ctx, c, f, 1e-3, 10, &p, bbox));
                p.passes_per_launch = 5;
                p.substeps_per_pass = 1;
                OK(st_evolve(ctx, c, f, 1e-3, 10, &p, nullptr));
                p.passes_per_launch = 1;
                double fr = 0;
                int64_t rs = 0;
                OK(st_cloud_stray_fraction(c, &fr));
                OK(st_cloud_resort_count(c, &rs));
                st_graph *g = nullptr;
                OK(st_graph_capture(ctx, c, f, 1e-3, 10, &p, 1, &g));
                for (int r = 0; r < 3; ++r) OK(st_graph_launch(g));
                expect(st_cloud_destroy(c), ST_E_STATE, "st_cloud_destroy of a cloud a graph holds");
                OK(st_graph_destroy(g));
            }
            // injection and, in the channel form, compaction
            Host<T> k;
            k.fill(700, m, 99 + bnd);
            OK(st_cloud_reserve(c, n + 2000));
            OK(st_cloud_inject(c, 700, k.d.data(), k.x.data(), k.y.data(), k.z.data(), k.u.data(), k.v.data(),
                               k.w.data(), nullptr));
            OK(st_evolve(ctx, c, f, 1e-3, 10, &p, nullptr));
            int64_t live = 0, cnt = 0;
            if (bnd == ST_CHANNEL) OK(st_cloud_compact(c, &live));
            OK(st_cloud_count(c, &cnt));
            Host<T> out;
            out.fill(cnt, m, 1);
            OK(st_cloud_download(c, out.d.data(), out.x.data(), out.y.data(), out.z.data(), out.u.data(),
                                 out.v.data(), out.w.data()));
            OK(st_field_merge_source(f));
            OK(st_field_download_source(f, src.data(), 8));
            std::vector<float> src32(ncell * 3);
            OK(st_field_download_source(f, src32.data(), 4));
            double ms = 0;
            int32_t launches = 0;
            OK(st_last_kernel_ms(ctx, &ms, &launches));
            OK(st_cloud_destroy(c));
            OK(st_field_destroy(f));
        }
}

void batch(st_ctx *ctx) {
    const int nb = 5;
    std::vector<st_cloud *> cs(nb);
    std::vector<st_field *> fs(nb);
    std::vector<Host<float>> hs(nb);
    for (int i = 0; i < nb; ++i) {
        const int N = 8 + 4 * i;
        const st_mesh m{{N, N, N}, {0, 0, 0}, {1, 1, 1}};
        std::vector<double> U = taylor_green(m);
        OK(st_field_create(ctx, &m, ST_FP32, &fs[i]));
        OK(st_field_upload_U(fs[i], U.data(), 8));
        OK(st_field_zero_source(fs[i]));
        const int64_t n = 300 + 1700 * i;
        OK(st_cloud_create(ctx, ST_FP32, n, 1e-3, 1000.0, 1.0, 0.0, 0, &cs[i]));
        hs[i].fill(n, m, 500 + i);
        OK(st_cloud_upload(cs[i], hs[i].d.data(), hs[i].x.data(), hs[i].y.data(), hs[i].z.data(),
                           hs[i].u.data(), hs[i].v.data(), hs[i].w.data(), nullptr));
    }
    st_params p;
    st_params_default(&p);
    p.boundary = ST_PERIODIC;
    p.substeps_per_pass = 1;
    st_tune_result tr;
    OK(st_autotune_batch(ctx, nb, cs.data(), fs.data(), &p, 1e-3, 1, nullptr, &tr));
    st_batch *b = nullptr;
    OK(st_batch_create(ctx, nb, cs.data(), fs.data(), &p, tr.target_parcels_per_brick, tr.blocks_per_cu, &b));
    OK(st_batch_evolve(b, 1e-3, 10));
    OK(st_batch_sort(b));
    int32_t items = 0, grid = 0;
    std::vector<int32_t> bricks(3 * nb);
    OK(st_batch_info(b, &items, &grid, bricks.data()));
    st_graph *g = nullptr;
    OK(st_batch_graph_capture(b, 1e-3, 10, 1, &g));
    OK(st_graph_launch(g));
    expect(st_batch_destroy(b), ST_E_STATE, "st_batch_destroy while its graph lives");
    OK(st_graph_destroy(g));
    OK(st_batch_destroy(b));
    for (int i = 0; i < nb; ++i) {
        OK(st_cloud_destroy(cs[i]));
        OK(st_field_destroy(fs[i]));
    }
}

void adjoint(st_ctx *ctx) {
    const st_mesh m{{12, 12, 12}, {0, 0, 0}, {1, 1, 1}};
    const int64_t n = 3000, nc = 12 * 12 * 12;
    std::vector<double> U = taylor_green(m);
    st_field *f = nullptr;
    st_cloud *c = nullptr;
    OK(st_field_create(ctx, &m, ST_FP64, &f));
    OK(st_field_upload_U(f, U.data(), 8));
    OK(st_cloud_create(ctx, ST_FP64, n, 1e-3, 0.0, 0.0, 1000.0, 1, &c));
    Host<double> h;
    h.fill(n, m, 4242);
    std::vector<double> rho(n, 2500.0);
    OK(st_cloud_upload(c, h.d.data(), h.x.data(), h.y.data(), h.z.data(), h.u.data(), h.v.data(), h.w.data(),
                       rho.data()));
    for (int bnd : {ST_PERIODIC, ST_WALL}) {
        st_params p;
        st_params_default(&p);
        p.boundary = bnd;
        st_adjoint *a = nullptr;
        OK(st_adjoint_create(ctx, c, f, 40, 10, &p, &a));
        OK(st_adjoint_forward(a, 4e-3));
        std::vector<double> target(nc * 3, 1e-12), xbar(3 * n, 0.5), vbar(3 * n, -0.25);
        double loss = 0;
        OK(st_adjoint_source_misfit(a, target.data(), &loss));
        OK(st_adjoint_set_state_cotangent(a, xbar.data(), vbar.data()));
        OK(st_adjoint_backward(a));
        std::vector<double> gd(n), grho(n), gx(3 * n), gv(3 * n);
        double gcd = 0;
        OK(st_adjoint_gradients(a, gd.data(), grho.data(), &gcd, gx.data(), gv.data()));
        OK(st_adjoint_set_source_cotangent(a, target.data()));
        OK(st_adjoint_backward(a));
        OK(st_adjoint_gradients(a, gd.data(), nullptr, &gcd, nullptr, nullptr));
        // page-locked host arrays for the target and the gradients, a fresh forward
        // with only the velocity cotangent seeded (the position one reads as zero)
        double *pt = nullptr, *pg = nullptr;
        OK(st_host_alloc((int64_t)(nc * 3 * sizeof(double)), (void **)&pt));
        OK(st_host_alloc((int64_t)(5 * n * sizeof(double)), (void **)&pg));  // d, rho, x0
        for (size_t k = 0; k < nc * 3; ++k) pt[k] = target[k];
        OK(st_adjoint_forward(a, 4e-3));
        OK(st_adjoint_source_misfit(a, pt, &loss));
        OK(st_adjoint_set_state_cotangent(a, nullptr, vbar.data()));
        OK(st_adjoint_backward(a));
        OK(st_adjoint_gradients(a, pg, pg + n, &gcd, pg + 2 * n, nullptr));
        for (int64_t k = 2 * n; k < 5 * n; ++k)
            if (pg[k] != 0.0) {  // dL/dx0 is the position cotangent, which was not given
                std::fprintf(stderr, "asan_driver: dL/dx0 not zero without a position cotangent\n");
                std::exit(1);
            }
        OK(st_host_free(pt));
        OK(st_host_free(pg));
        OK(st_adjoint_destroy(a));
    }
    OK(st_cloud_destroy(c));
    OK(st_field_destroy(f));
}

// An upload and an injection large enough that the library's host checks and staging
// copy run on several threads.
void threaded(st_ctx *ctx) {
    const st_mesh m{{32, 16, 16}, {0.0, 0.0, 0.0}, {2.0, 1.0, 1.0}};
    const int64_t n = 300000, k = 200000;
    std::vector<double> U = taylor_green(m);
    st_field *f = nullptr;
    st_cloud *c = nullptr;
    OK(st_field_create(ctx, &m, ST_FP32, &f));
    OK(st_field_upload_U(f, U.data(), 8));
    OK(st_cloud_create(ctx, ST_FP32, n, 1e-3, 1000.0, 1.0, 0.0, 0, &c));
    Host<float> h;
    h.fill(n, m, 31);
    OK(st_cloud_upload(c, h.d.data(), h.x.data(), h.y.data(), h.z.data(), h.u.data(), h.v.data(), h.w.data(),
                       nullptr));
    st_params p;
    st_params_default(&p);
    p.boundary = ST_CHANNEL;
    p.substeps_per_pass = 1;
    OK(st_field_zero_source(f));
    OK(st_evolve(ctx, c, f, 1e-3, 10, &p, nullptr));
    Host<float> inj;
    inj.fill(k, m, 32);
    OK(st_cloud_reserve(c, n + k));
    OK(st_cloud_inject(c, k, inj.d.data(), inj.x.data(), inj.y.data(), inj.z.data(), inj.u.data(), inj.v.data(),
                       inj.w.data(), nullptr));
    inj.x[k - 1] = 3.0f;  // outside the mesh: refused before anything is appended
    expect(st_cloud_inject(c, k, inj.d.data(), inj.x.data(), inj.y.data(), inj.z.data(), inj.u.data(), inj.v.data(),
                           inj.w.data(), nullptr),
           ST_E_RANGE, "st_cloud_inject with a parcel outside the mesh");
    OK(st_field_zero_source(f));
    OK(st_evolve(ctx, c, f, 1e-3, 10, &p, nullptr));
    OK(st_ctx_synchronize(ctx));
    OK(st_cloud_destroy(c));
    OK(st_field_destroy(f));
}

void refusals(st_ctx *ctx) {
    const st_mesh m{{6, 6, 6}, {0, 0, 0}, {1, 1, 1}};
    st_field *f = nullptr;
    st_cloud *c = nullptr;
    OK(st_field_create(ctx, &m, ST_FP64, &f));
    OK(st_cloud_create(ctx, ST_FP64, 100, 1e-3, 1000.0, 1.0, 0.0, 0, &c));
    expect(st_evolve(ctx, c, f, 1e-3, 10, nullptr, nullptr), ST_E_STATE, "st_evolve before upload");
    Host<double> h;
    h.fill(100, m, 3);
    h.x[10] = NAN;
    expect(st_cloud_upload(c, h.d.data(), h.x.data(), h.y.data(), h.z.data(), h.u.data(), h.v.data(), h.w.data(),
                           nullptr),
           ST_E_INVALID, "st_cloud_upload with a NaN");
    h.x[10] = 2.0;  // outside the mesh
    OK(st_cloud_upload(c, h.d.data(), h.x.data(), h.y.data(), h.z.data(), h.u.data(), h.v.data(), h.w.data(),
                       nullptr));
    std::vector<double> U = taylor_green(m);
    expect(st_field_upload_U(f, U.data(), 3), ST_E_INVALID, "st_field_upload_U with element size 3");
    OK(st_field_upload_U(f, U.data(), 8));
    expect(st_evolve(ctx, c, f, 1e-3, 10, nullptr, nullptr), ST_E_RANGE, "st_evolve with a parcel outside");
    expect(st_evolve(ctx, c, f, 1e-3, 0, nullptr, nullptr), ST_E_INVALID, "st_evolve with nsub 0");
    expect(st_evolve(ctx, c, f, NAN, 10, nullptr, nullptr), ST_E_INVALID, "st_evolve with dt NaN");
    expect(st_evolve(nullptr, c, f, 1e-3, 10, nullptr, nullptr), ST_E_INVALID, "st_evolve without context");
    st_params p;
    st_params_default(&p);
    p.block = 100;
    expect(st_evolve(ctx, c, f, 1e-3, 10, &p, nullptr), ST_E_INVALID, "st_evolve with block 100");
    expect(st_cloud_inject(c, 5, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
           ST_E_INVALID, "st_cloud_inject with null arrays");
    void *hp = nullptr;
    expect(st_host_alloc(0, &hp), ST_E_INVALID, "st_host_alloc of 0 bytes");
    OK(st_host_free(nullptr));
    OK(st_cloud_destroy(c));
    OK(st_field_destroy(f));
}

}  // namespace

int main() {
    int32_t devs = 0;
    OK(st_device_count(&devs));
    if (devs < 1) {
        std::fprintf(stderr, "asan_driver: no GPU\n");
        return 1;
    }
    if (st_version() != ST_ABI_VERSION) {
        std::fprintf(stderr, "asan_driver: ABI version %d, header %d\n", st_version(), ST_ABI_VERSION);
        return 1;
    }
    st_ctx *ctx = nullptr;
    OK(st_ctx_create(0, &ctx));
    double gbs = 0;
    OK(st_diag_copy_bandwidth(ctx, 1 << 24, 2, &gbs));
    double ms[8];
    int32_t launches[8], count = 0;
    OK(st_diag_kernel_times(ctx, 8, ms, launches, &count));
    forms<double>(ctx, ST_FP64);
    forms<float>(ctx, ST_FP32);
    batch(ctx);
    adjoint(ctx);
    threaded(ctx);
    refusals(ctx);
    // a destroyed context stays alive behind a cloud made on it, and refuses work
    st_ctx *c2 = nullptr;
    OK(st_ctx_create(0, &c2));
    st_cloud *cl = nullptr;
    OK(st_cloud_create(c2, ST_FP32, 10, 1e-3, 1000.0, 1.0, 0.0, 0, &cl));
    OK(st_ctx_destroy(c2));
    expect(st_ctx_synchronize(c2), ST_E_STATE, "st_ctx_synchronize after destroy");
    OK(st_cloud_destroy(cl));
    OK(st_ctx_synchronize(ctx));
    OK(st_ctx_destroy(ctx));
    std::printf("asan_driver: ok\n");
    return g_fail;
}
