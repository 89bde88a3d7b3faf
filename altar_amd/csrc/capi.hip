// C ABI of the altar_amd HIP library: source handles, priors, Metropolis and
// COV entry points.  Declarations and reference citations: include/altar_amd.h.
#include "../../include/altar_amd.h"

#include "common.hpp"
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#define ALTAR_TIMER_SLOTS 4

namespace {

thread_local std::string g_error;

struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw Failure(ALTAR_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

void require(bool ok, const std::string &msg)
{
    if (!ok) throw Failure(ALTAR_EINVAL, msg);
}

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return ALTAR_OK;
    } catch (const Failure &e) {
        g_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_error = "out of host memory";
        return ALTAR_ENOMEM;
    } catch (const std::exception &e) {
        g_error = e.what();
        return ALTAR_EHIP;
    } catch (...) {
        g_error = "unknown failure";
        return ALTAR_EHIP;
    }
}

// device buffer owned by a handle
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void reserve(size_t n, bool zero = false)
    {
        if (n <= bytes) return;
        release();
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            throw Failure(ALTAR_ENOMEM, std::string("hipMalloc(") + std::to_string(n) + "): " + hipGetErrorString(e));
        }
        bytes = n;
        if (zero) hip_check(hipMemset(p, 0, n), "hipMemset");
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return (T *)p; }
};

template <typename T>
void upload(DevBuf &b, const std::vector<T> &v)
{
    b.reserve(v.size() * sizeof(T));
    hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

inline hipStream_t S(void *s) { return (hipStream_t)s; }

__hip_bfloat16 to_bf16(float x) { return __float2bfloat16(x); }
float from_bf16(__hip_bfloat16 x) { return __bfloat162float(x); }

}  // namespace

using namespace altar;

struct altar_source {
    int model = 0, P = 0, N = 0;
    double nu = 0.25;
    int layout[8] = {0};
    int nsets = 0;
    bool have_layout = false, have_obs = false, have_green = false, prewhitened = false;
    size_t Npad = 0, Ppad = 0;
    // observations
    DevBuf x, y, lx, ly, lz, d, wd, xd, yd, lxd, lyd, lzd, dd, wdd, oid, zf, zd;
    // linear operator: G (Npad x Ppad FP32 and split BF16), d (Npad FP32)
    DevBuf G32, Ghi, Glo, dlin;
    // whitening
    int wkind = 0;
    double normalization = 0.0;
    DevBuf W32, Whi, Wlo, W64;
    // workspace, sized for Ccap chains (Cpad rows)
    int Ccap = 0;
    size_t Cpad = 0;
    DevBuf R32, Rhi, Rlo, R64, partial, T32, Thi, Tlo, prm, g32, g64, flags, shifts;
    // optional per-kernel timing with HIP events on the launch stream
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> marks[ALTAR_TIMER_SLOTS];

    ObsView obs(bool forward_only) const
    {
        ObsView o;
        o.N = N;
        o.x = x.as<float>(); o.y = y.as<float>();
        o.lx = lx.as<float>(); o.ly = ly.as<float>(); o.lz = lz.as<float>();
        o.d = forward_only ? zf.as<float>() : d.as<float>();
        o.wd = wd.as<float>();
        o.xd = xd.as<double>(); o.yd = yd.as<double>();
        o.lxd = lxd.as<double>(); o.lyd = lyd.as<double>(); o.lzd = lzd.as<double>();
        o.dd = forward_only ? zd.as<double>() : dd.as<double>();
        o.wdd = wdd.as<double>();
        o.oid = oid.as<int>();
        return o;
    }

    void ensure(int C)
    {
        if (C <= Ccap) return;
        Ccap = C;
        Cpad = round_up((size_t)C, 128);
        // zero-initialised: the GEMMs read the padding columns
        R32.release(); Rhi.release(); Rlo.release(); R64.release(); T32.release(); Thi.release(); Tlo.release();
        partial.release(); prm.release(); g32.release(); g64.release(); flags.release(); shifts.release();
    }
    size_t ntiles_obs() const { return (N + 255) / 256; }
};

struct altar_prior {
    int P = 0;
    DevBuf kind, a, b;
};

extern "C" {

const char *altar_last_error(void) { return g_error.c_str(); }

int altar_version(void) { return 100; }

// ------------------------------------------------------------------ sources
int altar_source_create(altar_source **out, int model, int P, int N, double nu)
{
    return guarded([&] {
        require(out != nullptr, "null handle pointer");
        require(model == ALTAR_MOGI || model == ALTAR_CDM || model == ALTAR_LINEAR, "unknown model");
        require(P > 0 && N > 0, "P and N must be positive");
        altar_source *h = new altar_source();
        h->model = model;
        h->P = P;
        h->N = N;
        h->nu = nu;
        h->Npad = round_up((size_t)N, 128);
        h->Ppad = round_up((size_t)P, 32);
        *out = h;
    });
}

int altar_source_destroy(altar_source *h)
{
    return guarded([&] {
        if (!h) return;
        DevBuf *all[] = {&h->x, &h->y, &h->lx, &h->ly, &h->lz, &h->d, &h->wd, &h->xd, &h->yd, &h->lxd,
                         &h->lyd, &h->lzd, &h->dd, &h->wdd, &h->oid, &h->zf, &h->zd, &h->G32, &h->Ghi,
                         &h->Glo, &h->dlin, &h->W32, &h->Whi, &h->Wlo, &h->W64, &h->R32, &h->Rhi,
                         &h->Rlo, &h->R64, &h->partial, &h->T32, &h->Thi, &h->Tlo, &h->prm, &h->g32,
                         &h->g64, &h->flags, &h->shifts};
        for (DevBuf *b : all) b->release();
        delete h;
    });
}

int altar_source_layout(altar_source *h, const int *idx, int nidx)
{
    return guarded([&] {
        require(h && idx, "null argument");
        int want = h->model == ALTAR_MOGI ? 5 : h->model == ALTAR_CDM ? 7 : 0;
        require(h->model != ALTAR_LINEAR, "the linear model has no layout");
        require(nidx == want, "layout: wrong number of indices");
        int nsets = idx[nidx - 1], off = idx[nidx - 2];
        require(nsets >= 0 && nsets <= MAX_SETS, "layout: nsets out of range [0, 8]");
        require((off < 0) == (nsets == 0), "layout: offsetIdx < 0 exactly when nsets == 0");
        for (int i = 0; i < nidx - 1; ++i)
            require(idx[i] < h->P, "layout: index beyond the parameter count");
        if (h->model == ALTAR_MOGI) {
            require(idx[0] >= 0 && idx[0] + 1 < h->P && idx[1] >= 0 && idx[2] >= 0, "layout: bad Mogi index");
        } else {
            require(idx[0] >= 0 && idx[0] + 1 < h->P && idx[1] >= 0 && idx[2] >= 0 && idx[3] >= 0 &&
                        idx[3] + 2 < h->P && idx[4] >= 0 && idx[4] + 2 < h->P,
                    "layout: bad CDM index");
        }
        if (nsets > 0) require(off + nsets <= h->P, "layout: offsets beyond the parameter count");
        for (int i = 0; i < nidx; ++i) h->layout[i] = idx[i];
        h->nsets = nsets;
        h->have_layout = true;
    });
}

int altar_source_observations(altar_source *h, const double *x, const double *y, const double *los,
                              const double *d, const int *oid)
{
    return guarded([&] {
        require(h && x && y && los && d && oid, "null argument");
        const int N = h->N;
        std::vector<float> fx(N), fy(N), flx(N), fly(N), flz(N), fd(N), zeros(N, 0.0f);
        std::vector<double> dlx(N), dly(N), dlz(N), zd(N, 0.0);
        std::vector<int> id(N);
        for (int n = 0; n < N; ++n) {
            require(oid[n] >= 0 && (h->model == ALTAR_LINEAR || oid[n] < std::max(h->nsets, 1)),
                    "observations: dataset id out of range");
            fx[n] = (float)x[n]; fy[n] = (float)y[n];
            dlx[n] = los[3 * n]; dly[n] = los[3 * n + 1]; dlz[n] = los[3 * n + 2];
            flx[n] = (float)dlx[n]; fly[n] = (float)dly[n]; flz[n] = (float)dlz[n];
            fd[n] = (float)d[n];
            id[n] = oid[n];
        }
        upload(h->x, fx); upload(h->y, fy);
        upload(h->lx, flx); upload(h->ly, fly); upload(h->lz, flz);
        upload(h->d, fd); upload(h->zf, zeros);
        upload(h->xd, std::vector<double>(x, x + N)); upload(h->yd, std::vector<double>(y, y + N));
        upload(h->lxd, dlx); upload(h->lyd, dly); upload(h->lzd, dlz);
        upload(h->dd, std::vector<double>(d, d + N)); upload(h->zd, zd);
        upload(h->oid, id);
        h->have_obs = true;
    });
}

int altar_source_green(altar_source *h, const double *G, const double *d, int prewhitened)
{
    return guarded([&] {
        require(h && G && d, "null argument");
        require(h->model == ALTAR_LINEAR, "green: not a linear source");
        const size_t N = h->N, P = h->P, Np = h->Npad, Pp = h->Ppad;
        std::vector<float> g(Np * Pp, 0.0f), dl(Np, 0.0f);
        std::vector<__hip_bfloat16> gh(Np * Pp, to_bf16(0.0f)), gl(Np * Pp, to_bf16(0.0f));
        for (size_t n = 0; n < N; ++n) {
            for (size_t p = 0; p < P; ++p) {
                float v = (float)G[n * P + p];
                g[n * Pp + p] = v;
                __hip_bfloat16 hi = to_bf16(v);
                gh[n * Pp + p] = hi;
                gl[n * Pp + p] = to_bf16(v - from_bf16(hi));
            }
            dl[n] = (float)d[n];
        }
        upload(h->G32, g); upload(h->Ghi, gh); upload(h->Glo, gl); upload(h->dlin, dl);
        // FP64 copies for the reference-precision path
        upload(h->dd, std::vector<double>(d, d + N));
        h->prewhitened = prewhitened != 0;
        h->have_green = true;
        h->have_obs = true;
    });
}

int altar_source_whitening(altar_source *h, int kind, const double *W, double normalization)
{
    return guarded([&] {
        require(h && W, "null argument");
        require(kind == ALTAR_WHITEN_DIAG || kind == ALTAR_WHITEN_LOWER, "whitening: unknown kind");
        const size_t N = h->N, Np = h->Npad;
        h->wkind = kind;
        h->normalization = normalization;
        if (kind == ALTAR_WHITEN_DIAG) {
            std::vector<float> w(N);
            for (size_t n = 0; n < N; ++n) w[n] = (float)W[n];
            upload(h->wd, w);
            upload(h->wdd, std::vector<double>(W, W + N));
            return;
        }
        std::vector<float> w(Np * Np, 0.0f);
        std::vector<__hip_bfloat16> wh(Np * Np, to_bf16(0.0f)), wl(Np * Np, to_bf16(0.0f));
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j <= i; ++j) {
                float v = (float)W[i * N + j];
                w[i * Np + j] = v;
                __hip_bfloat16 hi = to_bf16(v);
                wh[i * Np + j] = hi;
                wl[i * Np + j] = to_bf16(v - from_bf16(hi));
            }
        upload(h->W32, w); upload(h->Whi, wh); upload(h->Wlo, wl);
        std::vector<double> w64(N * N, 0.0);
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j <= i; ++j) w64[i * N + j] = W[i * N + j];
        upload(h->W64, w64);
    });
}

}  // extern "C"

namespace {

// brackets one kernel launch with events when the handle is timing
struct Mark {
    altar_source *h;
    int slot;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    Mark(altar_source *h_, int slot_, hipStream_t st_) : h(h_), slot(slot_), st(st_)
    {
        if (!h->timing) return;
        hip_check(hipEventCreate(&a), "hipEventCreate");
        hip_check(hipEventCreate(&b), "hipEventCreate");
        hip_check(hipEventRecord(a, st), "hipEventRecord");
    }
    ~Mark()
    {
        if (!h->timing || !a) return;
        if (hipEventRecord(b, st) == hipSuccess) h->marks[slot].push_back({a, b});
    }
};

void need_ready(altar_source *h, bool whitening)
{
    if (!h) throw Failure(ALTAR_EINVAL, "null handle");
    if (h->model != ALTAR_LINEAR && !h->have_layout) throw Failure(ALTAR_ESTATE, "layout not set");
    if (h->model == ALTAR_LINEAR ? !h->have_green : !h->have_obs)
        throw Failure(ALTAR_ESTATE, "observations not set");
    if (whitening && h->wkind == 0 && !(h->model == ALTAR_LINEAR && h->prewhitened))
        throw Failure(ALTAR_ESTATE, "whitening operator not set");
}

MogiArgs mogi_args(altar_source *h, const double *theta, int C, bool fwd)
{
    MogiArgs a;
    a.theta = theta; a.C = C; a.P = h->P;
    a.xIdx = h->layout[0]; a.dIdx = h->layout[1]; a.sIdx = h->layout[2]; a.offsetIdx = h->layout[3];
    a.nsets = h->nsets; a.nu = h->nu;
    h->prm.reserve((size_t)h->Ccap * MOGI_PRM * sizeof(double));
    a.prm = h->prm.as<double>();
    a.obs = h->obs(fwd);
    a.out = Out{};
    return a;
}

CdmArgs cdm_args(altar_source *h, const double *theta, int C, bool fwd)
{
    CdmArgs a;
    a.theta = theta; a.C = C; a.P = h->P;
    a.L = CdmLayout{h->layout[0], h->layout[1], h->layout[2], h->layout[3], h->layout[4], h->layout[5],
                    h->nsets};
    a.nu = h->nu;
    a.fp64_sin = 0.4;
    h->g32.reserve((size_t)h->Ccap * sizeof(cdm::Geom<float>));
    h->g64.reserve((size_t)h->Ccap * sizeof(cdm::Geom<double>));
    h->flags.reserve((size_t)h->Ccap * sizeof(unsigned));
    h->shifts.reserve((size_t)h->Ccap * MAX_SETS * sizeof(float));
    a.g32 = h->g32.as<cdm::Geom<float>>();
    a.g64 = h->g64.as<cdm::Geom<double>>();
    a.flags = h->flags.as<unsigned>();
    a.shifts = h->shifts.as<float>();
    a.obs = h->obs(fwd);
    a.out = Out{};
    return a;
}

// forward (+ residual) of a point-source model into one of the epilogues
void run_source(altar_source *h, const double *theta, int C, int mode, bool fp64, bool fwd, const Out &out,
                hipStream_t st)
{
    if (h->model == ALTAR_MOGI) {
        MogiArgs a = mogi_args(h, theta, C, fwd);
        a.out = out;
        Mark m(h, 0, st);
        hip_check(launch_mogi(a, mode, fp64 ? 1 : 0, st), "mogi kernel");
    } else {
        require(!fp64, "the CDM kernels compute in FP32 with FP64 near-singular sides; ALTAR_FP64 is not offered");
        CdmArgs a = cdm_args(h, theta, C, fwd);
        a.out = out;
        hip_check(launch_cdm_prepare(a, st), "cdm prepare");
        Mark m(h, 0, st);
        hip_check(launch_cdm_main(a, mode, st), "cdm kernel");
    }
}

// theta (C x P FP64) -> padded FP32 / split BF16 operand for the linear GEMMs
__global__ void theta_pack_kernel(const double *__restrict__ theta, int C, int P, size_t ld,
                                  float *__restrict__ t32, __hip_bfloat16 *__restrict__ thi,
                                  __hip_bfloat16 *__restrict__ tlo)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    int c = (int)(idx / P), p = (int)(idx % P);
    float v = (float)theta[idx];
    t32[(size_t)c * ld + p] = v;
    __hip_bfloat16 hi = __float2bfloat16(v);
    thi[(size_t)c * ld + p] = hi;
    tlo[(size_t)c * ld + p] = __float2bfloat16(v - __bfloat162float(hi));
}

// FP32 residual matrix -> split BF16 (for the linear model's BF16 whitening)
__global__ void split_kernel(const float *__restrict__ r, size_t n, __hip_bfloat16 *__restrict__ hi,
                             __hip_bfloat16 *__restrict__ lo)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = r[i];
    __hip_bfloat16 h = __float2bfloat16(v);
    hi[i] = h;
    lo[i] = __float2bfloat16(v - __bfloat162float(h));
}

void linear_residual_f32(altar_source *h, const double *theta, int C, bool fwd, hipStream_t st)
{
    const size_t Pp = h->Ppad, Np = h->Npad, Cp = h->Cpad;
    h->T32.reserve(Cp * Pp * sizeof(float), true);
    h->Thi.reserve(Cp * Pp * sizeof(__hip_bfloat16), true);
    h->Tlo.reserve(Cp * Pp * sizeof(__hip_bfloat16), true);
    hipLaunchKernelGGL(theta_pack_kernel, dim3((unsigned)(((long)C * h->P + 255) / 256)), dim3(256), 0, st, theta, C,
                       h->P, Pp, h->T32.as<float>(), h->Thi.as<__hip_bfloat16>(), h->Tlo.as<__hip_bfloat16>());
    hip_check(hipGetLastError(), "theta pack");
    h->R32.reserve(Cp * Np * sizeof(float), true);
    GemmEpi ep{};
    ep.bias = fwd ? nullptr : h->dlin.as<float>();
    ep.out = h->R32.as<float>();
    ep.ldo = Np;
    Mark m(h, 0, st);
    hip_check(launch_gemm_f32(h->T32.as<float>(), Pp, h->G32.as<float>(), Pp, (int)Cp, (int)Np, (int)Pp, 0,
                              EPI_STORE, ep, st),
              "linear gemm");
}

}  // namespace

extern "C" {

int altar_source_verify(altar_source *h, const double *theta_dev, int C, int *mask_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && mask_dev && C > 0, "bad argument");
        if (h->model != ALTAR_CDM) return;
        require(h->have_layout, "layout not set");
        CdmLayout L{h->layout[0], h->layout[1], h->layout[2], h->layout[3], h->layout[4], h->layout[5], h->nsets};
        hip_check(launch_cdm_verify(theta_dev, C, h->P, L, mask_dev, S(stream)), "cdm verify");
    });
}

}  // extern "C"

namespace {

__global__ void widen_kernel(const float *__restrict__ r, size_t ldr, int C, int N, double *__restrict__ out,
                             size_t ldo)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * N) return;
    int c = (int)(idx / N), n = (int)(idx % N);
    out[(size_t)c * ldo + n] = (double)r[(size_t)c * ldr + n];
}

void forward_common(altar_source *h, const double *theta, int C, double *out, size_t ldo, int precision,
                    bool fwd, hipStream_t st)
{
    need_ready(h, false);
    require(theta && out && C > 0 && ldo >= (size_t)h->N, "bad argument");
    require(precision == ALTAR_FP64 || precision == ALTAR_FP32, "forward: precision must be FP64 or FP32");
    h->ensure(C);
    if (h->model == ALTAR_LINEAR) {
        if (precision == ALTAR_FP64) {
            throw Failure(ALTAR_EINVAL, "linear forward: FP64 is not offered (use FP32)");
        }
        linear_residual_f32(h, theta, C, fwd, st);
        hipLaunchKernelGGL(widen_kernel, dim3((unsigned)(((long)C * h->N + 255) / 256)), dim3(256), 0, st,
                           h->R32.as<float>(), h->Npad, C, h->N, out, ldo);
        hip_check(hipGetLastError(), "widen");
        return;
    }
    Out o{};
    o.ld = ldo;
    o.r64 = out;
    run_source(h, theta, C, MODE_RES_F64, precision == ALTAR_FP64, fwd, o, st);
}

}  // namespace

extern "C" {

int altar_source_forward(altar_source *h, const double *theta_dev, int C, double *pred_dev, size_t ldo,
                         int precision, void *stream)
{
    return guarded([&] { forward_common(h, theta_dev, C, pred_dev, ldo, precision, true, S(stream)); });
}

int altar_source_residuals(altar_source *h, const double *theta_dev, int C, double *res_dev, size_t ldr,
                           int precision, void *stream)
{
    return guarded([&] { forward_common(h, theta_dev, C, res_dev, ldr, precision, false, S(stream)); });
}

int altar_source_data_llk(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision,
                          void *stream)
{
    return guarded([&] {
        need_ready(h, true);
        require(theta_dev && llk_dev && C > 0, "bad argument");
        require(precision >= ALTAR_FP64 && precision <= ALTAR_BF16_REFINED, "unknown precision");
        hipStream_t st = S(stream);
        h->ensure(C);
        const size_t Np = h->Npad, Cp = h->Cpad;
        const int ntn = (int)(Np / 128);

        if (h->model == ALTAR_LINEAR) {
            require(precision != ALTAR_FP64, "linear llk: FP64 is not offered");
            if (h->prewhitened) {
                // |G' theta - d'|^2 in one GEMM with the bias fused in the epilogue
                const size_t Pp = h->Ppad;
                h->T32.reserve(Cp * Pp * sizeof(float), true);
                h->Thi.reserve(Cp * Pp * sizeof(__hip_bfloat16), true);
                h->Tlo.reserve(Cp * Pp * sizeof(__hip_bfloat16), true);
                hipLaunchKernelGGL(theta_pack_kernel, dim3((unsigned)(((long)C * h->P + 255) / 256)), dim3(256), 0,
                                   st, theta_dev, C, h->P, Pp, h->T32.as<float>(), h->Thi.as<__hip_bfloat16>(),
                                   h->Tlo.as<__hip_bfloat16>());
                hip_check(hipGetLastError(), "theta pack");
                h->partial.reserve(Cp * ntn * sizeof(float));
                GemmEpi ep{};
                ep.bias = h->dlin.as<float>();
                ep.partial = h->partial.as<float>();
                ep.ntiles = ntn;
                Mark m(h, 1, st);
                if (precision == ALTAR_FP32)
                    hip_check(launch_gemm_f32(h->T32.as<float>(), Pp, h->G32.as<float>(), Pp, (int)Cp, (int)Np,
                                              (int)Pp, 0, EPI_SUMSQ_BIAS, ep, st),
                              "linear gemm");
                else
                    hip_check(launch_gemm_bf16x3(h->Thi.p, h->Tlo.p, Pp, h->Ghi.p, h->Glo.p, Pp, (int)Cp, (int)Np,
                                                 (int)Pp, 0, precision == ALTAR_BF16, EPI_SUMSQ_BIAS, ep, st),
                              "linear gemm");
                hip_check(launch_llk_reduce(h->partial.p, 0, ntn, C, h->normalization, llk_dev, st), "reduce");
                return;
            }
            linear_residual_f32(h, theta_dev, C, false, st);
        }

        if (h->wkind == ALTAR_WHITEN_DIAG) {
            if (h->model == ALTAR_LINEAR)
                throw Failure(ALTAR_EINVAL, "linear llk with a diagonal whitener: pre-whiten G and d instead");
            const int nt = (int)h->ntiles_obs();
            const bool fp64 = precision == ALTAR_FP64;
            h->partial.reserve(Cp * nt * sizeof(double));
            Out o{};
            o.partial = h->partial.p;
            run_source(h, theta_dev, C, MODE_DIAG, fp64, false, o, st);
            hip_check(launch_llk_reduce(h->partial.p, fp64 ? 1 : 0, nt, C, h->normalization, llk_dev, st),
                      "reduce");
            return;
        }

        // full lower-triangular whitening
        if (precision == ALTAR_FP64) {
            require(h->model == ALTAR_MOGI, "FP64 full whitening is offered for the Mogi source only");
            h->R64.reserve((size_t)C * h->N * sizeof(double));
            Out o{};
            o.ld = h->N;
            o.r64 = h->R64.as<double>();
            run_source(h, theta_dev, C, MODE_RES_F64, true, false, o, st);
            hip_check(launch_whiten_f64(h->R64.as<double>(), h->N, h->W64.as<double>(), h->N, h->N, C,
                                        h->normalization, llk_dev, st),
                      "whiten f64");
            return;
        }
        h->partial.reserve(Cp * ntn * sizeof(float));
        GemmEpi ep{};
        ep.partial = h->partial.as<float>();
        ep.ntiles = ntn;
        if (precision == ALTAR_FP32) {
            if (h->model != ALTAR_LINEAR) {
                h->R32.reserve(Cp * Np * sizeof(float), true);
                Out o{};
                o.ld = Np;
                o.r32 = h->R32.as<float>();
                run_source(h, theta_dev, C, MODE_RES_F32, false, false, o, st);
            }
            Mark m(h, 1, st);
            hip_check(launch_gemm_f32(h->R32.as<float>(), Np, h->W32.as<float>(), Np, (int)Cp, (int)Np, (int)Np, 1,
                                      EPI_SUMSQ, ep, st),
                      "whitening gemm");
        } else {
            h->Rhi.reserve(Cp * Np * sizeof(__hip_bfloat16), true);
            h->Rlo.reserve(Cp * Np * sizeof(__hip_bfloat16), true);
            if (h->model == ALTAR_LINEAR) {
                size_t n = Cp * Np;
                hipLaunchKernelGGL(split_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                                   h->R32.as<float>(), n, h->Rhi.as<__hip_bfloat16>(), h->Rlo.as<__hip_bfloat16>());
                hip_check(hipGetLastError(), "split");
            } else {
                Out o{};
                o.ld = Np;
                o.rhi = h->Rhi.as<__hip_bfloat16>();
                o.rlo = h->Rlo.as<__hip_bfloat16>();
                run_source(h, theta_dev, C, MODE_RES_BF16, false, false, o, st);
            }
            Mark m(h, 1, st);
            hip_check(launch_gemm_bf16x3(h->Rhi.p, h->Rlo.p, Np, h->Whi.p, h->Wlo.p, Np, (int)Cp, (int)Np, (int)Np, 1,
                                         precision == ALTAR_BF16, EPI_SUMSQ, ep, st),
                      "whitening gemm");
        }
        hip_check(launch_llk_reduce(h->partial.p, 0, ntn, C, h->normalization, llk_dev, st), "reduce");
    });
}

// ------------------------------------------------------------------ priors
int altar_prior_create(altar_prior **out, int P, const int *kind, const double *a, const double *b)
{
    return guarded([&] {
        require(out && kind && a && b && P > 0, "bad argument");
        for (int j = 0; j < P; ++j) {
            require(kind[j] == 0 || kind[j] == 1, "prior: unknown distribution kind");
            require(kind[j] == 0 ? b[j] > a[j] : b[j] > 0, "prior: empty support or non-positive sigma");
        }
        altar_prior *h = new altar_prior();
        h->P = P;
        upload(h->kind, std::vector<int>(kind, kind + P));
        upload(h->a, std::vector<double>(a, a + P));
        upload(h->b, std::vector<double>(b, b + P));
        *out = h;
    });
}

int altar_prior_destroy(altar_prior *h)
{
    return guarded([&] {
        if (!h) return;
        h->kind.release(); h->a.release(); h->b.release();
        delete h;
    });
}

int altar_prior_sample(altar_prior *h, double *theta_dev, int C, uint64_t seed, uint32_t stage, uint32_t attempt,
                       const int *only_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && C > 0, "bad argument");
        hip_check(launch_prior_sample(theta_dev, C, h->P, h->kind.as<int>(), h->a.as<double>(), h->b.as<double>(),
                                      seed, stage, attempt, only_dev, S(stream)),
                  "prior sample");
    });
}

int altar_prior_llk(altar_prior *h, const double *theta_dev, int C, double *llk_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && llk_dev && C > 0, "bad argument");
        hip_check(launch_prior_llk(theta_dev, C, h->P, h->kind.as<int>(), h->a.as<double>(), h->b.as<double>(),
                                   llk_dev, S(stream)),
                  "prior llk");
    });
}

int altar_prior_verify(altar_prior *h, const double *theta_dev, int C, int *mask_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && mask_dev && C > 0, "bad argument");
        hip_check(launch_prior_verify(theta_dev, C, h->P, h->kind.as<int>(), h->a.as<double>(), h->b.as<double>(),
                                      mask_dev, S(stream)),
                  "prior verify");
    });
}

// ------------------------------------------------------------------ Metropolis
int altar_mh_propose(const double *theta_dev, int C, int P, const double *L_dev, uint64_t seed, uint32_t stage,
                     uint32_t step, double *xi_dev, double *cand_dev, void *stream)
{
    return guarded([&] {
        require(theta_dev && L_dev && xi_dev && cand_dev && C > 0 && P > 0, "bad argument");
        hip_check(launch_normals(xi_dev, C, P, seed, stage, step, TAG_PROPOSAL, S(stream)), "normals");
        hip_check(launch_propose(theta_dev, L_dev, xi_dev, C, P, cand_dev, S(stream)), "propose");
    });
}

int altar_mh_restore(double *cand_dev, const double *theta_dev, const int *mask_dev, int C, int P, void *stream)
{
    return guarded([&] {
        require(cand_dev && theta_dev && mask_dev && C > 0 && P > 0, "bad argument");
        hip_check(launch_restore(cand_dev, theta_dev, mask_dev, C, P, S(stream)), "restore");
    });
}

int altar_mh_accept(double *theta_dev, double *prior_dev, double *data_dev, double *post_dev,
                    const double *ctheta_dev, const double *cprior_dev, const double *cdata_dev, double *cpost_dev,
                    const int *mask_dev, double beta, int C, int P, uint64_t seed, uint32_t stage, uint32_t step,
                    const double *dice_dev, int *decision_dev, unsigned long long *counters_dev, void *stream)
{
    return guarded([&] {
        require(theta_dev && prior_dev && data_dev && post_dev && ctheta_dev && cprior_dev && cdata_dev &&
                    mask_dev && counters_dev && C > 0 && P > 0,
                "bad argument");
        AcceptArgs a{theta_dev, prior_dev, data_dev, post_dev, ctheta_dev, cprior_dev, cdata_dev, cpost_dev,
                     mask_dev, beta, C, P, seed, stage, step, dice_dev, decision_dev, counters_dev};
        hip_check(launch_accept(a, S(stream)), "accept");
    });
}

double altar_mh_scaling(double aw, double rw, unsigned long long accepted, unsigned long long rejected,
                        unsigned long long unlikely)
{
    double total = (double)(accepted + rejected + unlikely);
    double acceptance = total > 0 ? (double)accepted / total : 0.0;
    double kc = (aw * acceptance + rw) / (aw + rw);
    if (kc < 0.1) kc = 0.1;
    if (kc > 1.0) kc = 1.0;
    return kc;
}

// ------------------------------------------------------------------ COV
}  // extern "C"

namespace {

// COV of the weights w = exp(dbeta (llk - median)) normalized to unit sum
// (cov::cov, COV.cc:334-368); returns the COV (1e100 if not finite)
double cov_of(double dbeta, const double *llk, int C, double median, std::vector<double> &w)
{
    double sum = 0.0;
    for (int i = 0; i < C; ++i) {
        w[i] = std::exp(dbeta * (llk[i] - median));
        sum += w[i];
    }
    const double wsum = C * (sum / C);
    const double scale = 1.0 / wsum;
    double msum = 0.0;
    for (int i = 0; i < C; ++i) {
        w[i] *= scale;
        msum += w[i];
    }
    const double mean = msum / C;
    double ss = 0.0;
    for (int i = 0; i < C; ++i) {
        const double dv = w[i] - mean;
        ss += dv * dv;
    }
    const double cov = std::sqrt(ss / (C - 1)) / mean;
    if (std::isinf(cov) || std::isnan(cov)) return 1e100;
    return cov;
}

}  // namespace

extern "C" {

int altar_cov_dbeta_grid(const double *llk, int C, double median, double beta, double tolerance, double target,
                         double *w_out, double *beta_out, double *cov_out)
{
    return guarded([&] {
        require(llk && w_out && beta_out && cov_out && C > 1, "bad argument");
        require(beta >= 0.0 && beta <= 1.0, "beta outside [0, 1]");
        std::vector<double> w(C);
        // COV.cc:222-326: a 10-interval grid refined around the first point
        // whose COV reaches the target
        double lo = 0.0, hi = 1.0 - beta;
        double cov = cov_of(hi, llk, C, median, w);
        double dbeta;
        if (cov < target || std::fabs(cov - target) < tolerance) {
            dbeta = hi;
            beta = 1.0;
        } else {
            const int Nb = 10;
            int nb = 0, loops = 0;
            bool found = false;
            double guess = lo, step;
            do {
                ++loops;
                step = (hi - lo) / Nb;
                for (guess = lo, nb = 0; nb <= Nb; guess += step, ++nb) {
                    cov = cov_of(guess, llk, C, median, w);
                    if (std::fabs(cov - target) < tolerance) { found = true; break; }
                    if (cov >= target) { if (nb == 0) found = true; break; }
                }
                if (nb > Nb) guess -= step;
                hi = guess;
                lo = guess - step;
            } while (loops <= Nb && !found);
            dbeta = guess;
            cov = cov_of(dbeta, llk, C, median, w);
            beta += dbeta;
        }
        std::memcpy(w_out, w.data(), sizeof(double) * C);
        *beta_out = beta;
        *cov_out = cov;
    });
}

int altar_cov_covariance(const double *theta_dev, const double *w_dev, int C, int P, double *mean_dev,
                         double *sigma_dev, void *stream)
{
    return guarded([&] {
        require(theta_dev && w_dev && mean_dev && sigma_dev && C > 0 && P > 0, "bad argument");
        hip_check(launch_wcov(theta_dev, w_dev, C, P, mean_dev, sigma_dev, S(stream)), "covariance");
    });
}

int altar_cov_multiplicities(const double *w, int C, uint64_t seed, uint32_t stage, int *count)
{
    return guarded([&] {
        require(w && count && C > 0, "bad argument");
        std::vector<double> edge(C + 1);
        edge[0] = 0.0;
        for (int i = 0; i < C; ++i) edge[i + 1] = edge[i] + w[i];
        edge[C] = 1.0;
        std::fill(count, count + C, 0);
        for (int i = 0; i < C; ++i) {
            u32x4 r = draw(seed, stage, 0u, (uint32_t)i, TAG_RESAMPLE, 0u);
            const double u = u53(r.x, r.y);
            // bins are (edge[k], edge[k+1]]: smallest k with u <= edge[k+1]
            const int k = (int)(std::lower_bound(edge.begin() + 1, edge.end(), u) - (edge.begin() + 1));
            count[std::min(k, C - 1)] += 1;
        }
    });
}

int altar_resample_gather(const double *theta_old, const double *prior_old, const double *data_old,
                          const double *post_old, const int *src_dev, int C, int P, double *theta, double *prior,
                          double *data, double *post, void *stream)
{
    return guarded([&] {
        require(theta_old && prior_old && data_old && post_old && src_dev && theta && prior && data && post &&
                    C > 0 && P > 0,
                "bad argument");
        hip_check(launch_gather(theta_old, prior_old, data_old, post_old, src_dev, C, P, theta, prior, data, post,
                                S(stream)),
                  "gather");
    });
}

int altar_posterior(const double *prior_dev, const double *data_dev, double beta, int C, double *post_dev,
                    void *stream)
{
    return guarded([&] {
        require(prior_dev && data_dev && post_dev && C > 0, "bad argument");
        hip_check(launch_posterior(prior_dev, data_dev, beta, C, post_dev, S(stream)), "posterior");
    });
}

int altar_random_normals(double *out_dev, int C, int P, uint64_t seed, uint32_t stage, uint32_t step, uint32_t tag,
                         void *stream)
{
    return guarded([&] {
        require(out_dev && C > 0 && P > 0, "bad argument");
        hip_check(launch_normals(out_dev, C, P, seed, stage, step, tag, S(stream)), "normals");
    });
}

}  // extern "C"

extern "C" {

int altar_source_timing(altar_source *h, int enable)
{
    return guarded([&] {
        require(h != nullptr, "null handle");
        h->timing = enable != 0;
    });
}

int altar_source_timers(altar_source *h, double *ms, long long *count, int nslots)
{
    return guarded([&] {
        require(h && ms && count && nslots > 0, "bad argument");
        for (int s = 0; s < ALTAR_TIMER_SLOTS; ++s) {
            double tot = 0.0;
            for (auto &p : h->marks[s]) {
                hip_check(hipEventSynchronize(p.second), "hipEventSynchronize");
                float t = 0.0f;
                hip_check(hipEventElapsedTime(&t, p.first, p.second), "hipEventElapsedTime");
                tot += t;
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
            if (s < nslots) {
                ms[s] = tot;
                count[s] = (long long)h->marks[s].size();
            }
            h->marks[s].clear();
        }
    });
}

}  // extern "C"
