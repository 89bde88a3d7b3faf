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
#include <mutex>
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

// one diagonal block of a block-diagonal whitener: observations [off, off + n)
// (off a multiple of 128), lower-triangular W of n x n stored npad x npad
struct WBlock {
    int off = 0, n = 0, npad = 0;
    DevBuf W32, Whi, Wlo;
};

struct altar_source {
    int model = 0, P = 0, N = 0;
    double nu = 0.25;
    double fp64_sin = 0.2;  // CDM sides with |sin beta| below this run in FP64
    int layout[10] = {0};
    int nsets = 0;
    bool have_layout = false, have_obs = false, have_green = false, prewhitened = false;
    size_t Npad = 0, Ppad = 0;
    // observations
    DevBuf x, y, lx, ly, lz, d, wd, xd, yd, lxd, lyd, lzd, dd, wdd, oid, zf, zd;
    // linear operator: G (Npad x Ppad FP32 and split BF16), d (Npad FP32)
    DevBuf G32, Ghi, Glo, G64, cbias, cref;
    // whitening
    int wkind = 0;
    double normalization = 0.0;
    DevBuf W32, Whi, Wlo, W64;
    std::vector<WBlock> blocks;  // ALTAR_WHITEN_BLOCKS
    // workspace, sized for Ccap chains (Cpad rows)
    int Ccap = 0;
    size_t Cpad = 0;
    DevBuf R32, Rhi, Rlo, R64, partial, T32, Thi, Tlo, prm, g32, g64, flags, shifts, census;
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
        R32.release(); Rhi.release(); Rlo.release(); R64.release(); T32.release(); Thi.release();
        Tlo.release();
        partial.release(); prm.release(); g32.release(); g64.release(); flags.release(); shifts.release(); census.release();
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
        require(model == ALTAR_MOGI || model == ALTAR_CDM || model == ALTAR_LINEAR || model == ALTAR_JOINT,
                "unknown model");
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
                         &h->lyd, &h->lzd, &h->dd, &h->wdd, &h->oid, &h->zf, &h->zd, &h->G64, &h->cbias, &h->cref, &h->G32, &h->Ghi,
                         &h->Glo, &h->W32, &h->Whi, &h->Wlo, &h->W64, &h->R32, &h->Rhi,
                         &h->Rlo, &h->R64, &h->partial, &h->T32, &h->Thi, &h->Tlo, &h->prm, &h->g32,
                         &h->g64, &h->flags, &h->shifts};
        for (DevBuf *b : all) b->release();
        for (auto &slot : h->marks)
            for (auto &p : slot) {
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
        h->census.release();
        for (WBlock &b : h->blocks) {
            b.W32.release(); b.Whi.release(); b.Wlo.release();
        }
        delete h;
    });
}

int altar_source_layout(altar_source *h, const int *idx, int nidx)
{
    return guarded([&] {
        require(h && idx, "null argument");
        int want = h->model == ALTAR_MOGI ? 5 : h->model == ALTAR_CDM ? 7 : h->model == ALTAR_JOINT ? 10 : 0;
        require(h->model != ALTAR_LINEAR, "the linear model has no layout");
        require(nidx == want, "layout: wrong number of indices");
        int nsets = idx[nidx - 1], off = idx[nidx - 2];
        require(nsets >= 0 && nsets <= MAX_SETS, "layout: nsets out of range [0, 8]");
        require((off < 0) == (nsets == 0), "layout: offsetIdx < 0 exactly when nsets == 0");
        for (int i = 0; i < nidx - 1; ++i)
            require(idx[i] < h->P, "layout: index beyond the parameter count");
        auto mogi_ok = [&](const int *q) { return q[0] >= 0 && q[0] + 1 < h->P && q[1] >= 0 && q[2] >= 0; };
        auto cdm_ok = [&](const int *q) {
            return q[0] >= 0 && q[0] + 1 < h->P && q[1] >= 0 && q[2] >= 0 && q[3] >= 0 && q[3] + 2 < h->P &&
                   q[4] >= 0 && q[4] + 2 < h->P;
        };
        if (h->model == ALTAR_MOGI) {
            require(mogi_ok(idx), "layout: bad Mogi index");
        } else if (h->model == ALTAR_CDM) {
            require(cdm_ok(idx), "layout: bad CDM index");
        } else {
            require(mogi_ok(idx) && cdm_ok(idx + 3), "layout: bad joint index");
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
        std::vector<float> g(Np * Pp, 0.0f);
        std::vector<__hip_bfloat16> gh(Np * Pp, to_bf16(0.0f)), gl(Np * Pp, to_bf16(0.0f));
        for (size_t n = 0; n < N; ++n) {
            for (size_t p = 0; p < P; ++p) {
                float v = (float)G[n * P + p];
                g[n * Pp + p] = v;
                __hip_bfloat16 hi = to_bf16(v);
                gh[n * Pp + p] = hi;
                gl[n * Pp + p] = to_bf16(v - from_bf16(hi));
            }
        }
        upload(h->G32, g); upload(h->Ghi, gh); upload(h->Glo, gl);
        // FP64 copies: the centred residual (linear_center) and the
        // reference-precision path
        upload(h->G64, std::vector<double>(G, G + N * P));
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
        for (WBlock &B : h->blocks) {
            B.W32.release(); B.Whi.release(); B.Wlo.release();
        }
        h->blocks.clear();
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
        std::vector<double> w64(Np * Np, 0.0);
        for (size_t i = 0; i < N; ++i)
            for (size_t j = 0; j <= i; ++j) w64[i * Np + j] = W[i * N + j];
        upload(h->W64, w64);
    });
}

int altar_source_whitening_blocks(altar_source *h, int nblocks, const int *offsets, const int *sizes,
                                  const double *W, double normalization)
{
    return guarded([&] {
        require(h && offsets && sizes && W && nblocks > 0, "null argument");
        int end = 0;
        for (int b = 0; b < nblocks; ++b) {
            require(offsets[b] % 128 == 0, "whitening blocks: offsets must be multiples of 128");
            require(sizes[b] > 0 && offsets[b] >= end && offsets[b] + sizes[b] <= h->N,
                    "whitening blocks: blocks must be ordered, disjoint and inside [0, N)");
            end = offsets[b] + sizes[b];
        }
        std::vector<WBlock> blocks(nblocks);
        auto release = [](std::vector<WBlock> &v) {
            for (WBlock &B : v) {
                B.W32.release(); B.Whi.release(); B.Wlo.release();
            }
        };
        try {
            size_t at = 0;
            for (int b = 0; b < nblocks; ++b) {
                WBlock &B = blocks[b];
                B.off = offsets[b];
                B.n = sizes[b];
                B.npad = (int)round_up((size_t)B.n, 128);
                const size_t n = B.n, np = B.npad;
                std::vector<float> w(np * np, 0.0f);
                std::vector<__hip_bfloat16> wh(np * np, to_bf16(0.0f)), wl(np * np, to_bf16(0.0f));
                for (size_t i = 0; i < n; ++i)
                    for (size_t j = 0; j <= i; ++j) {
                        float v = (float)W[at + i * n + j];
                        w[i * np + j] = v;
                        __hip_bfloat16 hi = to_bf16(v);
                        wh[i * np + j] = hi;
                        wl[i * np + j] = to_bf16(v - from_bf16(hi));
                    }
                upload(B.W32, w); upload(B.Whi, wh); upload(B.Wlo, wl);
                at += n * n;
            }
        } catch (...) {
            release(blocks);
            throw;
        }
        release(h->blocks);
        h->blocks = std::move(blocks);
        h->wkind = ALTAR_WHITEN_BLOCKS;
        h->normalization = normalization;
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
    a.xIdx = h->layout[0]; a.dIdx = h->layout[1]; a.sIdx = h->layout[2];
    a.offsetIdx = h->model == ALTAR_JOINT ? h->layout[8] : h->layout[3];
    a.nsets = h->nsets; a.nu = h->nu;
    h->prm.reserve((size_t)h->Ccap * MOGI_PRM * sizeof(double));
    a.prm = h->prm.as<double>();
    a.obs = h->obs(fwd);
    a.out = Out{};
    return a;
}

// the CDM part of a CDM or joint layout; in a joint source the dataset
// shifts belong to the Mogi part (applied once)
CdmLayout cdm_layout(const altar_source *h)
{
    if (h->model == ALTAR_JOINT)
        return CdmLayout{h->layout[3], h->layout[4], h->layout[5], h->layout[6], h->layout[7], -1, 0};
    return CdmLayout{h->layout[0], h->layout[1], h->layout[2], h->layout[3], h->layout[4], h->layout[5], h->nsets};
}

CdmArgs cdm_args(altar_source *h, const double *theta, int C, bool fwd)
{
    CdmArgs a;
    a.theta = theta; a.C = C; a.P = h->P;
    a.L = cdm_layout(h);
    a.nu = h->nu;
    a.fp64_sin = h->fp64_sin;
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
    } else if (h->model == ALTAR_JOINT) {
        // superposition: the Mogi kernel writes pred - shift - d, the CDM
        // kernel adds its prediction
        require(!fp64, "the joint source computes in FP32 (FP64 near-singular CDM sides); ALTAR_FP64 is not offered");
        require(mode == MODE_RES_F32 || mode == MODE_RES_F64, "joint source: residual modes only");
        MogiArgs m = mogi_args(h, theta, C, fwd);
        m.out = out;
        {
            Mark mk(h, 2, st);
            hip_check(launch_mogi(m, mode, 0, st), "mogi kernel");
        }
        CdmArgs a = cdm_args(h, theta, C, fwd);
        a.out = out;
        hip_check(launch_cdm_prepare(a, st), "cdm prepare");
        Mark mk(h, 0, st);
        hip_check(launch_cdm_main(a, mode == MODE_RES_F32 ? MODE_ACC_F32 : MODE_ACC_F64, st), "cdm kernel");
    } else {
        require(!fp64, "the CDM kernels compute in FP32 with FP64 near-singular sides; ALTAR_FP64 is not offered");
        CdmArgs a = cdm_args(h, theta, C, fwd);
        a.out = out;
        hip_check(launch_cdm_prepare(a, st), "cdm prepare");
        Mark m(h, 0, st);
        hip_check(launch_cdm_main(a, mode, st), "cdm kernel");
    }
}

// The linear residual is formed about a reference sample theta_ref (the
// mean of the call's first 64 chains): G theta - d = G (theta - theta_ref) - b with
// b = d - G theta_ref in FP64.  The GEMM then multiplies small offsets, and
// the cancellation between G theta and d (whitened data are tens of times
// the residual) happens in FP64 instead of in the FP32/BF16 product.
// theta_ref[p] = mean of theta[0 .. min(C, 64) - 1][p], in chain order
constexpr int REF_ROWS = 64;
__global__ void theta_ref_kernel(const double *__restrict__ theta, int C, int P, double *__restrict__ ref)
{
    const int n = min(C, REF_ROWS);
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        double s = 0.0;
#pragma unroll 16
        for (int c = 0; c < n; ++c) s += theta[(size_t)c * P + p];
        ref[p] = s / n;
    }
}

// b[n] for n < N, one wavefront per observation; zero past N.
__global__ void center_bias_kernel(const double *__restrict__ G, const double *__restrict__ d,
                                   const double *__restrict__ theta_ref, int N, int Np, int P,
                                   float *__restrict__ b)
{
    const int n = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (n >= Np) return;
    double acc = 0.0;
    if (n < N)
        for (int p = lane; p < P; p += 64) acc += G[(size_t)n * P + p] * theta_ref[p];
    acc = wave_sum(acc);
    if (lane == 0) b[n] = n < N ? (float)(d[n] - acc) : 0.0f;
}

// theta (C x P FP64) -> padded FP32 / split BF16 operand for the linear
// GEMMs, as offsets from theta_ref
__global__ void theta_pack_kernel(const double *__restrict__ theta, int C, int P, size_t ld,
                                  const double *__restrict__ theta_ref, float *__restrict__ t32,
                                  __hip_bfloat16 *__restrict__ thi, __hip_bfloat16 *__restrict__ tlo)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    int c = (int)(idx / P), p = (int)(idx % P);
    float v = (float)(theta[idx] - theta_ref[p]);
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

// b = d - G theta_ref (centred residual) into h->cbias, and the packed
// offsets theta - theta_ref
void linear_center(altar_source *h, const double *theta, int C, hipStream_t st)
{
    const size_t Pp = h->Ppad, Np = h->Npad, Cp = h->Cpad;
    h->cref.reserve((size_t)h->P * sizeof(double));
    hipLaunchKernelGGL(theta_ref_kernel, dim3(1), dim3(256), 0, st, theta, C, h->P, h->cref.as<double>());
    hip_check(hipGetLastError(), "centre");
    const double *ref = h->cref.as<double>();
    h->cbias.reserve(Np * sizeof(float), true);
    hipLaunchKernelGGL(center_bias_kernel, dim3((unsigned)((Np * 64 + 255) / 256)), dim3(256), 0, st,
                       h->G64.as<double>(), h->dd.as<double>(), ref, h->N, (int)Np, h->P, h->cbias.as<float>());
    hip_check(hipGetLastError(), "centre bias");
    h->T32.reserve(Cp * Pp * sizeof(float), true);
    h->Thi.reserve(Cp * Pp * sizeof(__hip_bfloat16), true);
    h->Tlo.reserve(Cp * Pp * sizeof(__hip_bfloat16), true);
    hipLaunchKernelGGL(theta_pack_kernel, dim3((unsigned)(((long)C * h->P + 255) / 256)), dim3(256), 0, st, theta, C,
                       h->P, Pp, ref, h->T32.as<float>(), h->Thi.as<__hip_bfloat16>(), h->Tlo.as<__hip_bfloat16>());
    hip_check(hipGetLastError(), "theta pack");
}

// theta (C x P FP64) -> padded FP32 operand without centring (forward)
__global__ void theta_pack_plain_kernel(const double *__restrict__ theta, int C, int P, size_t ld,
                                        float *__restrict__ t32)
{
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    int c = (int)(idx / P), p = (int)(idx % P);
    t32[(size_t)c * ld + p] = (float)theta[idx];
}

void linear_residual_f32(altar_source *h, const double *theta, int C, bool fwd, hipStream_t st)
{
    // buffers are sized for the handle's capacity; the launches cover the
    // rows of this call only (C rounded up to the 128-row tile)
    const size_t Pp = h->Ppad, Np = h->Npad, Cp = h->Cpad, Mp = round_up((size_t)C, 128);
    if (fwd) {
        // predictions G theta: no centring
        h->T32.reserve(Cp * Pp * sizeof(float), true);
        hipLaunchKernelGGL(theta_pack_plain_kernel, dim3((unsigned)(((long)C * h->P + 255) / 256)), dim3(256), 0,
                           st, theta, C, h->P, Pp, h->T32.as<float>());
        hip_check(hipGetLastError(), "theta pack");
    } else {
        linear_center(h, theta, C, st);
    }
    h->R32.reserve(Cp * Np * sizeof(float), true);
    GemmEpi ep{};
    ep.bias = fwd ? nullptr : h->cbias.as<float>();
    ep.out = h->R32.as<float>();
    ep.ldo = Np;
    Mark m(h, 0, st);
    hip_check(launch_gemm_f32(h->T32.as<float>(), Pp, h->G32.as<float>(), Pp, (int)Mp, (int)Np, (int)Pp, 0,
                              EPI_STORE, ep, st),
              "linear gemm");
}

}  // namespace

extern "C" {

int altar_source_verify(altar_source *h, const double *theta_dev, int C, int *mask_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && mask_dev && C > 0, "bad argument");
        if (h->model != ALTAR_CDM && h->model != ALTAR_JOINT) return;
        require(h->have_layout, "layout not set");
        hip_check(launch_cdm_verify(theta_dev, C, h->P, cdm_layout(h), mask_dev, S(stream)), "cdm verify");
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

namespace {
void data_llk_impl(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision, float *wr_out,
                   size_t ldw, hipStream_t st);
}

int altar_source_data_llk(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision,
                          void *stream)
{
    return guarded([&] { data_llk_impl(h, theta_dev, C, llk_dev, precision, nullptr, 0, S(stream)); });
}

int altar_source_data_llk_fields(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision,
                                 float *wr_dev, size_t ldw, void *stream)
{
    return guarded([&] {
        require(h != nullptr, "null handle");
        require(wr_dev != nullptr && ldw >= h->Npad, "fields: output row pitch must be >= the padded N");
        require(h->wkind != 0 || (h->model == ALTAR_LINEAR && h->prewhitened), "fields: whitening operator not set");
        require(precision == ALTAR_FP32 || precision == ALTAR_BF16 || precision == ALTAR_BF16_REFINED,
                "fields: precision must be FP32 or BF16 (the stored fields are FP32)");
        data_llk_impl(h, theta_dev, C, llk_dev, precision, wr_dev, ldw, S(stream));
    });
}

}  // extern "C"

namespace {
// Whitened fields of a diagonal whitener (or of a pre-whitened linear
// residual, w = NULL): block (c, t) writes z = w r for the 128 columns
// [128 t, 128 t + 128) of chain c (zero past N) and the sum of z^2 into
// partial[c ntiles + t], the layout llk_reduce_kernel sums in tile order.
__global__ __launch_bounds__(128) void diag_fields_kernel(const float *__restrict__ R, size_t ldr,
                                                          const float *__restrict__ w, int N,
                                                          float *__restrict__ out, size_t ldo,
                                                          float *__restrict__ partial, int ntiles)
{
    const int c = blockIdx.x, t = blockIdx.y;
    const int n = t * 128 + threadIdx.x;
    float z = 0.0f;
    if (n < N) {
        const float r = R[(size_t)c * ldr + n];
        z = w ? w[n] * r : r;
    }
    out[(size_t)c * ldo + n] = z;
    __shared__ float half[2];
    const float q = wave_sum(z * z);
    if ((threadIdx.x & 63) == 0) half[threadIdx.x >> 6] = q;
    __syncthreads();
    if (threadIdx.x == 0) partial[(size_t)c * ntiles + t] = half[0] + half[1];
}

void data_llk_impl(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision, float *wr_out,
                   size_t ldw, hipStream_t stream)
{
    {
        need_ready(h, true);
        require(theta_dev && llk_dev && C > 0, "bad argument");
        require(precision >= ALTAR_FP64 && precision <= ALTAR_BF16_REFINED, "unknown precision");
        hipStream_t st = stream;
        h->ensure(C);
        // buffers are sized for the handle's capacity (Cp rows); every launch
        // covers only the rows of this call, C rounded up to the 128-row tile
        // (an output buffer of the caller has exactly those rows)
        const size_t Np = h->Npad, Cp = h->Cpad, Mp = round_up((size_t)C, 128);
        const int ntn = (int)(Np / 128);

        if (wr_out && (h->wkind == ALTAR_WHITEN_DIAG || (h->model == ALTAR_LINEAR && h->prewhitened))) {
            // fields of a diagonal whitener or of a pre-whitened linear
            // source: FP32 residuals, then z = w r per element with its
            // per-tile sums of squares (no GEMM); FP32 whatever the
            // whitening precision asked for
            require(h->model != ALTAR_JOINT, "joint fields need a lower or block whitener");
            require(h->model != ALTAR_LINEAR || h->prewhitened,
                    "linear fields with a diagonal whitener: pre-whiten G and d instead");
            h->R32.reserve(Cp * Np * sizeof(float), true);
            if (h->model == ALTAR_LINEAR) {
                linear_residual_f32(h, theta_dev, C, false, st);
            } else {
                Out o{};
                o.ld = Np;
                o.r32 = h->R32.as<float>();
                run_source(h, theta_dev, C, MODE_RES_F32, false, false, o, st);
            }
            h->partial.reserve(Cp * ntn * sizeof(float));
            const float *w = h->model == ALTAR_LINEAR ? nullptr : h->wd.as<float>();
            hipLaunchKernelGGL(diag_fields_kernel, dim3((unsigned)C, (unsigned)ntn), dim3(128), 0, st,
                               h->R32.as<float>(), Np, w, h->N, wr_out, ldw, h->partial.as<float>(), ntn);
            hip_check(hipGetLastError(), "diagonal fields");
            hip_check(launch_llk_reduce(h->partial.p, 0, ntn, C, h->normalization, llk_dev, st), "reduce");
            return;
        }
        if (h->model == ALTAR_LINEAR) {
            require(precision != ALTAR_FP64, "linear llk: FP64 is not offered");
            if (h->prewhitened) {
                // |G' theta - d'|^2 in one GEMM with the (centred) bias fused in the epilogue
                const size_t Pp = h->Ppad;
                linear_center(h, theta_dev, C, st);
                h->partial.reserve(Cp * ntn * sizeof(float));
                GemmEpi ep{};
                ep.bias = h->cbias.as<float>();
                ep.partial = h->partial.as<float>();
                ep.ntiles = ntn;
                Mark m(h, 1, st);
                if (precision == ALTAR_FP32)
                    hip_check(launch_gemm_f32(h->T32.as<float>(), Pp, h->G32.as<float>(), Pp, (int)Mp, (int)Np,
                                              (int)Pp, 0, EPI_SUMSQ_BIAS, ep, st),
                              "linear gemm");
                else
                    hip_check(launch_gemm_bf16x3(h->Thi.p, h->Tlo.p, Pp, h->Ghi.p, h->Glo.p, Pp, (int)Mp, (int)Np,
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
            if (h->model == ALTAR_JOINT)
                throw Failure(ALTAR_EINVAL, "joint llk with a diagonal whitener: pass it as a lower or block whitener");
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

        // full lower-triangular whitening, or block-diagonal (one triangular
        // GEMM per block over the block's columns of R)
        const bool blocked = h->wkind == ALTAR_WHITEN_BLOCKS;
        if (precision == ALTAR_FP64) {
            // the reference's precision: FP64 residuals (Mogi in FP64; CDM
            // with every side in FP64), |W r|^2 by the FP64 MFMA GEMM with
            // its sum-of-squares epilogue, FP64 sums
            require(!blocked && (h->model == ALTAR_MOGI || h->model == ALTAR_CDM),
                    "FP64 full whitening is offered for Mogi and CDM sources with a lower-triangular whitener");
            require(h->model != ALTAR_CDM || h->fp64_sin > 1.0,
                    "CDM FP64 llk: evaluate every side in FP64 first (altar_source_fp64_sides(h, 2))");
            // padding columns stay zero (the source writes columns below N only)
            h->R64.reserve(Cp * Np * sizeof(double), true);
            Out o{};
            o.ld = Np;
            o.r64 = h->R64.as<double>();
            run_source(h, theta_dev, C, MODE_RES_F64, h->model == ALTAR_MOGI, false, o, st);
            h->partial.reserve(Cp * ntn * sizeof(double));
            {
                Mark m(h, 1, st);
                hip_check(launch_gemm_f64_sumsq(h->R64.as<double>(), Np, h->W64.as<double>(), Np, (int)Mp, (int)Np,
                                                (int)Np, 1, h->N, h->partial.as<double>(), ntn, st),
                          "FP64 whitening");
            }
            hip_check(launch_llk_reduce(h->partial.p, 1, ntn, C, h->normalization, llk_dev, st), "reduce");
            return;
        }
        int ntiles = ntn;
        if (blocked) {
            ntiles = 0;
            int at = 0;
            for (const WBlock &B : h->blocks) {
                ntiles += B.npad / 128;
                // fields: the blocks' padded column ranges must tile [0, Npad), so that
                // every column of the caller's buffer is written
                require(!wr_out || B.off == at, "fields: the whitening blocks must tile the padded observations");
                at = B.off + B.npad;
            }
            require(!wr_out || at == (int)Np, "fields: the whitening blocks must tile the padded observations");
        }
        h->partial.reserve(Cp * ntiles * sizeof(float));
        GemmEpi ep{};
        ep.partial = h->partial.as<float>();
        ep.ntiles = ntiles;
        int epi = EPI_SUMSQ;
        if (wr_out) {
            // the whitened residuals are stored too; rows past C land in a
            // workspace (the caller's buffer has C rows)
            require(C % 128 == 0, "fields: C must be a multiple of 128");
            ep.out = wr_out;
            ep.ldo = ldw;
            epi = EPI_STORE_SUMSQ;
        }
        const bool bf16 = precision != ALTAR_FP32;
        // residual matrix: FP32 (then split for BF16 when the source writes FP32 only)
        const bool via_f32 = !bf16 || h->model != ALTAR_MOGI;
        if (via_f32 && h->model != ALTAR_LINEAR) {
            h->R32.reserve(Cp * Np * sizeof(float), true);
            Out o{};
            o.ld = Np;
            o.r32 = h->R32.as<float>();
            run_source(h, theta_dev, C, MODE_RES_F32, false, false, o, st);
        }
        if (bf16) {
            h->Rhi.reserve(Cp * Np * sizeof(__hip_bfloat16), true);
            h->Rlo.reserve(Cp * Np * sizeof(__hip_bfloat16), true);
            if (via_f32) {
                size_t n = Mp * Np;
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
        }
        // one triangular GEMM: columns [off, off + npad) of R against W (npad x npad)
        // nvalid: the block's observations; W's rows past them are zero
        auto whiten = [&](size_t off, int npad, int nvalid, const DevBuf &W32, const DevBuf &Whi, const DevBuf &Wlo,
                          int tile0) {
            GemmEpi e = ep;
            e.partial = ep.partial + tile0;
            e.nvalid = nvalid;
            if (wr_out) e.out = wr_out + off;  // this block's columns of the whitened fields
            if (!bf16)
                hip_check(launch_gemm_f32(h->R32.as<float>() + off, Np, W32.as<float>(), (size_t)npad, (int)Mp, npad,
                                          npad, 1, epi, e, st),
                          "whitening gemm");
            else
                hip_check(launch_gemm_bf16x3(h->Rhi.as<__hip_bfloat16>() + off, h->Rlo.as<__hip_bfloat16>() + off, Np,
                                             Whi.p, Wlo.p, (size_t)npad, (int)Mp, npad, npad, 1,
                                             precision == ALTAR_BF16, epi, e, st),
                          "whitening gemm");
        };
        {
            Mark m(h, 1, st);
            if (!blocked) {
                whiten(0, (int)Np, h->N, h->W32, h->Whi, h->Wlo, 0);
            } else {
                int tile0 = 0;
                for (const WBlock &B : h->blocks) {
                    whiten((size_t)B.off, B.npad, B.n, B.W32, B.Whi, B.Wlo, tile0);
                    tile0 += B.npad / 128;
                }
            }
        }
        hip_check(launch_llk_reduce(h->partial.p, 0, ntiles, C, h->normalization, llk_dev, st), "reduce");
    }
}
}  // namespace

extern "C" {

// Y = A B^T (- bias per column), FP32 MFMA; M, N multiples of 128, K of 16;
// row pitches multiples of 4 elements
int altar_gemm_f32(const float *A_dev, size_t lda, const float *B_dev, size_t ldb, int M, int N, int K,
                   const float *bias_dev, float *out_dev, size_t ldo, void *stream)
{
    return guarded([&] {
        require(A_dev && B_dev && out_dev && M > 0 && N > 0 && K > 0, "bad argument");
        require(M % 128 == 0 && N % 128 == 0 && K % 16 == 0, "gemm: M, N must be multiples of 128 and K of 16");
        require(lda % 4 == 0 && ldb % 4 == 0 && lda >= (size_t)K && ldb >= (size_t)K && ldo >= (size_t)N,
                "gemm: bad row pitch");
        GemmEpi ep{};
        ep.bias = bias_dev;
        ep.out = out_dev;
        ep.ldo = ldo;
        hip_check(launch_gemm_f32(A_dev, lda, B_dev, ldb, M, N, K, 0, EPI_STORE, ep, S(stream)), "gemm");
    });
}

// dst[c, r] = src[r, c] (FP32, rows x cols)
int altar_transpose_f32(const float *src_dev, size_t lds, int rows, int cols, float *dst_dev, size_t ldd,
                        void *stream)
{
    return guarded([&] {
        require(src_dev && dst_dev && rows > 0 && cols > 0 && lds >= (size_t)cols && ldd >= (size_t)rows,
                "bad argument");
        hip_check(launch_transpose_f32(src_dev, lds, rows, cols, dst_dev, ldd, S(stream)), "transpose");
    });
}

// dst[r, :cols] = src[idx[r], :cols] for r < n (FP32 rows)
int altar_gather_rows_f32(const float *src_dev, size_t lds, const int *idx_dev, int n, int cols, float *dst_dev,
                          size_t ldd, void *stream)
{
    return guarded([&] {
        require(src_dev && idx_dev && dst_dev && n >= 0 && cols > 0, "bad argument");
        hip_check(launch_gather_rows_f32(src_dev, lds, idx_dev, n, cols, dst_dev, ldd, S(stream)), "gather rows");
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

int altar_cov_dbeta_brent(const double *llk, int C, double median, double beta, double tolerance, double target,
                          int maxiter, double *w_out, double *beta_out, double *cov_out)
{
    return guarded([&] {
        require(llk && w_out && beta_out && cov_out && C > 1 && maxiter >= 0, "bad argument");
        require(beta >= 0.0 && beta <= 1.0, "beta outside [0, 1]");
        std::vector<double> w(C);
        double last_cov = 0.0;
        // the minimized metric (COV - target)^2, 1e100 where the COV is not finite (COV.cc:355-362)
        auto metric = [&](double dbeta) {
            last_cov = cov_of(dbeta, llk, C, median, w);
            return last_cov == 1e100 ? 1e100 : (last_cov - target) * (last_cov - target);
        };
        double lo = 0.0, hi = 1.0 - beta;
        metric(hi);
        if (last_cov < target || std::fabs(last_cov - target) < tolerance) {   // skip to beta = 1
            std::memcpy(w_out, w.data(), sizeof(double) * C);
            *beta_out = 1.0;
            *cov_out = last_cov;
            return;
        }
        // Brent's method (Brent 1973, localmin): golden-section steps, parabolic
        // steps through the best three points when they fall inside the bracket
        const double golden = 0.5 * (3.0 - std::sqrt(5.0));
        const double rel_eps = 1.4901161193847656e-08, abs_eps = 1e-12;
        double best = 5.0e-5, f_best = metric(best);
        double second = best, third = best, f_second = f_best, f_third = f_best;
        double step = 0.0, prev_step = 0.0;
        for (int it = 0; it < maxiter && !(std::fabs(last_cov - target) < tolerance); ++it) {
            const double mid = 0.5 * (lo + hi);
            const double t1 = rel_eps * std::fabs(best) + abs_eps, t2 = 2.0 * t1;
            if (std::fabs(best - mid) <= t2 - 0.5 * (hi - lo)) break;
            double num = 0.0, den = 0.0, older = 0.0;
            if (std::fabs(prev_step) > t1) {
                const double r = (best - second) * (f_best - f_third);
                den = (best - third) * (f_best - f_second);
                num = (best - third) * den - (best - second) * r;
                den = 2.0 * (den - r);
                if (den > 0.0) num = -num; else den = -den;
                older = prev_step;
                prev_step = step;
            }
            if (std::fabs(num) < std::fabs(0.5 * den * older) && num > den * (lo - best) && num < den * (hi - best)) {
                step = num / den;
                const double trial = best + step;
                if (trial - lo < t2 || hi - trial < t2) step = best < mid ? t1 : -t1;
            } else {
                prev_step = best < mid ? hi - best : lo - best;
                step = golden * prev_step;
            }
            const double u = std::fabs(step) >= t1 ? best + step : best + (step > 0.0 ? t1 : -t1);
            const double fu = metric(u);
            if (fu <= f_best) {
                if (u < best) hi = best; else lo = best;
                third = second; f_third = f_second;
                second = best; f_second = f_best;
                best = u; f_best = fu;
            } else {
                if (u < best) lo = u; else hi = u;
                if (fu <= f_second || second == best) {
                    third = second; f_third = f_second;
                    second = u; f_second = fu;
                } else if (fu <= f_third || third == best || third == second) {
                    third = u; f_third = fu;
                }
            }
        }
        metric(best);
        std::memcpy(w_out, w.data(), sizeof(double) * C);
        *beta_out = beta + best;
        *cov_out = last_cov;
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

int altar_source_fp64_sides(altar_source *h, double threshold)
{
    return guarded([&] {
        require(h != nullptr, "null handle");
        require(h->model == ALTAR_CDM || h->model == ALTAR_JOINT, "fp64 sides: a CDM or joint source is required");
        require(threshold >= 0.0 && threshold <= 2.0, "fp64 sides: threshold must lie in [0, 2]");
        h->fp64_sin = threshold;
    });
}

int altar_source_side_census(altar_source *h, int C, long long *counts, void *stream)
{
    return guarded([&] {
        require(h != nullptr && counts != nullptr && C > 0, "bad argument");
        require(h->model == ALTAR_CDM || h->model == ALTAR_JOINT, "side census: a CDM or joint source is required");
        require(h->flags.p != nullptr && (size_t)C <= h->flags.bytes / sizeof(unsigned),
                "side census: no forward call has prepared this many chains");
        hipStream_t st = S(stream);
        h->census.reserve(4 * sizeof(unsigned long long));
        hip_check(launch_side_census(h->flags.as<unsigned>(), C, h->census.as<unsigned long long>(), st), "census");
        unsigned long long host[4];
        hip_check(hipMemcpyAsync(host, h->census.p, sizeof(host), hipMemcpyDeviceToHost, st), "census copy");
        hip_check(hipStreamSynchronize(st), "census sync");
        for (int i = 0; i < 4; ++i) counts[i] = (long long)host[i];
    });
}

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
        hipError_t first = hipSuccess;
        for (int s = 0; s < ALTAR_TIMER_SLOTS; ++s) {
            double tot = 0.0;
            for (auto &p : h->marks[s]) {
                float t = 0.0f;
                hipError_t e = hipEventSynchronize(p.second);
                if (e == hipSuccess) e = hipEventElapsedTime(&t, p.first, p.second);
                if (e != hipSuccess && first == hipSuccess) first = e;
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
        hip_check(first, "timing events");
    });
}

}  // extern "C"

// ------------------------------------------------------------------ surrogate MLP
// Dense network in -> h -> ... -> h -> out with GELU on the hidden layers,
// BF16 weights (FP32 master copies, Adam), FP32 accumulation.
struct altar_mlp {
    int nin = 0, nhid = 0, nlayers = 0, nout = 0, maxB = 0;
    int kin = 0, kout = 0;                  // padded input / output widths
    std::vector<int> din, dout;             // padded layer widths
    std::vector<DevBuf> W32, Wm, Wv, W16, WT16, b32, bm, bv, gW, gb;
    std::vector<DevBuf> pre, act16, actT16, dY16, dYT16;
    DevBuf x16, xT16, out32, lo, hi, loss, tgt, wk;
    DevBuf xb;                              // a training batch's inputs (altar_mlp_train_epoch)
    DevBuf adam_segs;                       // the Adam step's tensor table (W, b of each layer)
    int adam_blocks = 0;
    std::vector<DevBuf> cparts, gparts;     // per-layer column-sum and split-k partials of the backward pass
    DevBuf wgrad_tab, reduce_tab;           // segment tables of the grouped weight-gradient and sum launches
    int tab_rows = -1, wgrad_blocks = 0, reduce_blocks = 0;   // the batch rows the tables were built for
    long step = 0;
    int P = 0;
};

namespace {

// B input rows; the GEMMs run on B rounded up to the 64-row tile, Bp (input
// rows B..Bp-1 are zero).  With train set, the transposes of the layer
// inputs (Bp columns) are written beside them for the weight gradients.
void mlp_forward(altar_mlp *h, const double *theta, int B, hipStream_t st, bool train = false)
{
    const int Bp = (int)round_up((size_t)B, 64);
    hip_check(launch_mlp_input(theta, B, h->P, h->lo.as<double>(), h->hi.as<double>(), h->x16.p, h->kin, Bp,
                               train ? h->xT16.p : nullptr, Bp, st),
              "mlp input");
    B = Bp;
    const void *X = h->x16.p;
    int ldx = h->kin;
    for (int l = 0; l < h->nlayers; ++l) {
        MlpEpi ep{};
        ep.bias = h->b32[l].as<float>();
        ep.pre = h->pre[l].as<float>();
        ep.out16 = h->act16[l].as<__bf16>();
        ep.out32 = (l == h->nlayers - 1) ? h->out32.as<float>() : nullptr;
        ep.ldo = h->dout[l];
        ep.act = l < h->nlayers - 1;
        ep.out16T = (train && l < h->nlayers - 1) ? h->actT16[l].as<__bf16>() : nullptr;
        ep.ldT = Bp;
        hip_check(launch_mlp_gemm(MLP_EPI_BIAS_ACT, X, ldx, h->W16[l].p, h->din[l], B, h->dout[l], h->din[l], ep, st),
                  "mlp gemm");
        X = h->act16[l].p;
        ldx = h->dout[l];
    }
}

// k-ranges of the weight-gradient GEMM over a batch of B rows (B a multiple
// of 64): the most slices, up to 16, that keep each range >= 128 rows and a
// multiple of the 32-row k tile
constexpr int SPLITK_MAX = 16;
int splitk_slices(int B)
{
    for (int S = SPLITK_MAX; S > 1; S /= 2)
        if (B % (S * 32) == 0 && B / S >= 128) return S;
    return 1;
}

// Segment tables of the grouped weight-gradient GEMM and ordered sums for a
// batch of B rows (B a multiple of 64); rebuilt when B changes, after the
// stream has drained, since kernels of an earlier batch may still read them.
void mlp_backward_tables(altar_mlp *h, int B, hipStream_t st)
{
    if (h->tab_rows == B) return;
    hip_check(hipStreamSynchronize(st), "stream sync");
    const int L = h->nlayers, S = splitk_slices(B);
    std::vector<MlpWGradSeg> wg;
    std::vector<MlpReduceSeg> rd;
    int wb = 0, rb = 0;
    for (int l = L - 1; l >= 0; --l) {
        const int O = h->dout[l], K = h->din[l];
        const __bf16 *XT = l == 0 ? h->xT16.as<__bf16>() : h->actT16[l - 1].as<__bf16>();
        wg.push_back({h->dYT16[l].as<__bf16>(), XT, h->gparts[l].as<float>(), O, K, wb});
        wb += (O / 64) * (K / 64) * S;
        rd.push_back({h->gparts[l].as<float>(), h->gW[l].as<float>(), (long)O * K, (size_t)O * K, S, rb});
        rb += (O * K + 255) / 256;
        rd.push_back({h->cparts[l].as<float>(), h->gb[l].as<float>(), (long)O, (size_t)O, colsum_slabs(B), rb});
        rb += (O + 255) / 256;
    }
    upload(h->wgrad_tab, wg);
    upload(h->reduce_tab, rd);
    h->wgrad_blocks = wb;
    h->reduce_blocks = rb;
    h->tab_rows = B;
}

// Backward pass over B rows (a multiple of 64) after mlp_forward with train
// set.  Each dY is written with its transpose and the slab partials of its
// column sums by its producer: the loss gradient for the last layer, the
// backward-data GEMM of layer l + 1 for layer l.  Then one launch forms
// every layer's split-k weight-gradient partials and one adds them, and the
// column-sum partials, in order.  loss (optional) accumulates the batch's
// loss; it must start at zero.
void mlp_backward(altar_mlp *h, const float *target, const float *w, int B, int valid, float *loss, hipStream_t st)
{
    const int L = h->nlayers;
    mlp_backward_tables(h, B, st);
    hip_check(launch_mse_grad(h->out32.as<float>(), target, w, B, h->dout[L - 1], h->dout[L - 1], valid,
                              h->dY16[L - 1].p, h->dYT16[L - 1].p, B, h->cparts[L - 1].as<float>(), loss, st),
              "mse grad");
    for (int l = L - 1; l >= 1; --l) {
        const int O = h->dout[l], K = h->din[l];
        // data gradient through the previous GELU: dX = (dY . W) * gelu'(pre)
        MlpEpi ed{};
        ed.pre = h->pre[l - 1].as<float>();
        ed.out16 = h->dY16[l - 1].as<__bf16>();
        ed.ldo = K;
        ed.out16T = h->dYT16[l - 1].as<__bf16>();
        ed.ldT = B;
        ed.cpart = h->cparts[l - 1].as<float>();
        hip_check(launch_mlp_gemm(MLP_EPI_DGELU, h->dY16[l].p, O, h->WT16[l].p, O, B, K, O, ed, st), "mlp dX");
    }
    hip_check(launch_mlp_wgrad(h->wgrad_tab.as<MlpWGradSeg>(), L, h->wgrad_blocks, B, splitk_slices(B), B, st),
              "mlp dW");
    hip_check(launch_mlp_reduce(h->reduce_tab.as<MlpReduceSeg>(), 2 * L, h->reduce_blocks, st), "mlp sums");
}

// Gradients of the weighted squared error over a batch of B rows (no
// update), with the batch's targets and the output weights already in tgt
// and wk.  Columns nout..kout-1 of both are zero from their allocation and
// nothing writes them; rows B.. of tgt are never read.
void mlp_staged_gradients(altar_mlp *h, const double *theta_dev, int B, float *loss_dev, hipStream_t st)
{
    const int Bp = (int)round_up((size_t)B, 64);
    // rows B..Bp-1 of the input are zero and their targets are ignored
    mlp_forward(h, theta_dev, B, st, true);
    if (loss_dev) hip_check(hipMemsetAsync(h->loss.p, 0, 4, st), "memset");
    mlp_backward(h, h->tgt.as<float>(), h->wk.as<float>(), Bp, B, loss_dev ? h->loss.as<float>() : nullptr, st);
    if (loss_dev) hip_check(hipMemcpyAsync(loss_dev, h->loss.p, 4, hipMemcpyDeviceToDevice, st), "loss");
}

// the same from caller-held targets (B x nout, row pitch ldt) and weights
void mlp_batch_gradients(altar_mlp *h, const double *theta_dev, const float *target_dev, int ldt, const float *w_dev,
                         int B, float *loss_dev, hipStream_t st)
{
    hip_check(hipMemcpy2DAsync(h->tgt.p, h->kout * 4, target_dev, ldt * 4, h->nout * 4, B,
                               hipMemcpyDeviceToDevice, st),
              "targets");
    hip_check(hipMemcpyAsync(h->wk.p, w_dev, h->nout * 4, hipMemcpyDeviceToDevice, st), "weights");
    mlp_staged_gradients(h, theta_dev, B, loss_dev, st);
}

void mlp_adam_step(altar_mlp *h, float lr, float beta1, float beta2, float eps, hipStream_t st)
{
    h->step += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)h->step), c2 = 1.0f - std::pow(beta2, (float)h->step);
    hip_check(launch_adam(h->adam_segs.as<AdamSeg>(), 2 * h->nlayers, h->adam_blocks, lr, beta1, beta2, eps, c1, c2,
                          st),
              "adam");
}

}  // namespace

extern "C" {

int altar_mlp_create(altar_mlp **out, int nin, int nhid, int nlayers, int nout, int max_batch)
{
    return guarded([&] {
        require(out && nin > 0 && nhid > 0 && nlayers >= 2 && nout > 0 && max_batch > 0, "bad argument");
        require(nhid % 64 == 0, "hidden width must be a multiple of 64");
        altar_mlp *h = new altar_mlp();
        h->nin = nin; h->nhid = nhid; h->nlayers = nlayers; h->nout = nout;
        h->maxB = (int)round_up((size_t)max_batch, 64);
        h->kin = (int)round_up((size_t)nin, 64);
        h->kout = (int)round_up((size_t)nout, 64);
        h->P = nin;
        for (int l = 0; l < nlayers; ++l) {
            h->din.push_back(l == 0 ? h->kin : nhid);
            h->dout.push_back(l == nlayers - 1 ? h->kout : nhid);
        }
        const int L = nlayers;
        h->W32.resize(L); h->Wm.resize(L); h->Wv.resize(L); h->W16.resize(L); h->WT16.resize(L);
        h->b32.resize(L); h->bm.resize(L); h->bv.resize(L); h->gW.resize(L); h->gb.resize(L);
        h->pre.resize(L); h->act16.resize(L); h->actT16.resize(L); h->dY16.resize(L); h->dYT16.resize(L);
        const size_t B = h->maxB;
        for (int l = 0; l < L; ++l) {
            const size_t O = h->dout[l], K = h->din[l];
            h->W32[l].reserve(O * K * 4, true); h->Wm[l].reserve(O * K * 4, true); h->Wv[l].reserve(O * K * 4, true);
            h->W16[l].reserve(O * K * 2, true); h->WT16[l].reserve(O * K * 2, true);
            h->b32[l].reserve(O * 4, true); h->bm[l].reserve(O * 4, true); h->bv[l].reserve(O * 4, true);
            h->gW[l].reserve(O * K * 4, true); h->gb[l].reserve(O * 4, true);
            h->pre[l].reserve(B * O * 4, true); h->act16[l].reserve(B * O * 2, true);
            h->actT16[l].reserve(B * O * 2, true);
            h->dY16[l].reserve(B * O * 2, true); h->dYT16[l].reserve(B * O * 2, true);
        }
        h->x16.reserve(B * h->kin * 2, true);
        h->xT16.reserve(B * h->kin * 2, true);
        h->out32.reserve(B * h->kout * 4, true);
        h->lo.reserve(nin * 8, true);
        h->hi.reserve(nin * 8, true);
        h->loss.reserve(4, true);
        h->tgt.reserve(B * h->kout * 4, true);
        h->wk.reserve(h->kout * 4, true);
        h->cparts.resize(L);
        h->gparts.resize(L);
        for (int l = 0; l < L; ++l) {
            h->cparts[l].reserve((size_t)colsum_slabs((int)B) * h->dout[l] * 4, true);
            h->gparts[l].reserve((size_t)SPLITK_MAX * h->dout[l] * h->din[l] * 4, true);
        }
        std::vector<AdamSeg> segs;
        int blocks = 0;
        for (int l = 0; l < L; ++l) {
            const int O = h->dout[l], K = h->din[l];
            segs.push_back({h->W32[l].as<float>(), h->Wm[l].as<float>(), h->Wv[l].as<float>(), h->gW[l].as<float>(),
                            h->W16[l].as<__bf16>(), h->WT16[l].as<__bf16>(), O, K, blocks});
            blocks += (O * K + 255) / 256;
            segs.push_back({h->b32[l].as<float>(), h->bm[l].as<float>(), h->bv[l].as<float>(), h->gb[l].as<float>(),
                            nullptr, nullptr, 1, O, blocks});
            blocks += (O + 255) / 256;
        }
        upload(h->adam_segs, segs);
        h->adam_blocks = blocks;
        *out = h;
    });
}

int altar_mlp_destroy(altar_mlp *h)
{
    return guarded([&] {
        if (!h) return;
        for (auto *v : {&h->W32, &h->Wm, &h->Wv, &h->W16, &h->WT16, &h->b32, &h->bm, &h->bv, &h->gW, &h->gb,
                        &h->pre, &h->act16, &h->actT16, &h->dY16, &h->dYT16, &h->cparts, &h->gparts})
            for (auto &b : *v) b.release();
        for (DevBuf *b : {&h->x16, &h->xT16, &h->out32, &h->lo, &h->hi, &h->loss, &h->tgt, &h->wk, &h->xb,
                          &h->adam_segs, &h->wgrad_tab, &h->reduce_tab})
            b->release();
        delete h;
    });
}

int altar_mlp_input_scaling(altar_mlp *h, const double *lo, const double *hi)
{
    return guarded([&] {
        require(h && lo && hi, "bad argument");
        for (int j = 0; j < h->nin; ++j) require(hi[j] > lo[j], "input scaling: empty range");
        hip_check(hipMemcpy(h->lo.p, lo, h->nin * 8, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(h->hi.p, hi, h->nin * 8, hipMemcpyHostToDevice), "H2D");
    });
}

// layer l weights (out x in, row-major, unpadded) and biases (out), FP32 host
int altar_mlp_set_layer(altar_mlp *h, int l, const float *W, const float *b)
{
    return guarded([&] {
        require(h && W && b && l >= 0 && l < h->nlayers, "bad argument");
        const int O = h->dout[l], K = h->din[l];
        const int o_real = l == h->nlayers - 1 ? h->nout : h->nhid;
        const int k_real = l == 0 ? h->nin : h->nhid;
        std::vector<float> w(O * K, 0.0f), bb(O, 0.0f);
        std::vector<__hip_bfloat16> w16(O * K, to_bf16(0.0f)), wt16(O * K, to_bf16(0.0f));
        for (int o = 0; o < o_real; ++o) {
            for (int k = 0; k < k_real; ++k) {
                w[o * K + k] = W[o * k_real + k];
                w16[o * K + k] = to_bf16(W[o * k_real + k]);
                wt16[k * O + o] = to_bf16(W[o * k_real + k]);
            }
            bb[o] = b[o];
        }
        hip_check(hipMemcpy(h->W32[l].p, w.data(), w.size() * 4, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(h->W16[l].p, w16.data(), w16.size() * 2, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(h->WT16[l].p, wt16.data(), wt16.size() * 2, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(h->b32[l].p, bb.data(), bb.size() * 4, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemset(h->Wm[l].p, 0, w.size() * 4), "memset");
        hip_check(hipMemset(h->Wv[l].p, 0, w.size() * 4), "memset");
        hip_check(hipMemset(h->bm[l].p, 0, bb.size() * 4), "memset");
        hip_check(hipMemset(h->bv[l].p, 0, bb.size() * 4), "memset");
        h->step = 0;
    });
}

// read back layer l: FP32 master weights and biases (unpadded); with grads != 0
// the last computed gradients instead
int altar_mlp_get_layer(altar_mlp *h, int l, float *W, float *b, int grads)
{
    return guarded([&] {
        require(h && W && b && l >= 0 && l < h->nlayers, "bad argument");
        hip_check(hipDeviceSynchronize(), "sync");
        const int O = h->dout[l], K = h->din[l];
        const int o_real = l == h->nlayers - 1 ? h->nout : h->nhid;
        const int k_real = l == 0 ? h->nin : h->nhid;
        std::vector<float> w(O * K), bb(O);
        hip_check(hipMemcpy(w.data(), grads ? h->gW[l].p : h->W32[l].p, w.size() * 4, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(bb.data(), grads ? h->gb[l].p : h->b32[l].p, bb.size() * 4, hipMemcpyDeviceToHost), "D2H");
        for (int o = 0; o < o_real; ++o) {
            for (int k = 0; k < k_real; ++k) W[o * k_real + k] = w[o * K + k];
            b[o] = bb[o];
        }
    });
}

// z_dev (B x ldz, FP32) = network(theta_dev); B <= max_batch
int altar_mlp_forward(altar_mlp *h, const double *theta_dev, int B, float *z_dev, int ldz, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && z_dev && B > 0 && B <= h->maxB && ldz >= h->nout, "bad argument");
        hipStream_t st = S(stream);
        mlp_forward(h, theta_dev, B, st);
        hip_check(hipMemcpy2DAsync(z_dev, ldz * 4, h->out32.p, h->kout * 4, h->nout * 4, B, hipMemcpyDeviceToDevice, st),
                  "copy out");
    });
}

// gradients of the weighted squared error over a batch (no update):
// loss = sum_b sum_k w_k (net(theta_b)_k - target_bk)^2 / B; target is B x nout
// (FP32, device, row pitch ldt); loss_dev (optional) receives the loss
int altar_mlp_backward(altar_mlp *h, const double *theta_dev, const float *target_dev, int ldt, const float *w_dev,
                       int B, float *loss_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_dev && target_dev && w_dev && B > 0 && B <= h->maxB, "bad argument");
        mlp_batch_gradients(h, theta_dev, target_dev, ldt, w_dev, B, loss_dev, S(stream));
    });
}

// one Adam step on the gradients of the last altar_mlp_backward
int altar_mlp_adam(altar_mlp *h, float lr, float beta1, float beta2, float eps, void *stream)
{
    return guarded([&] {
        require(h && lr > 0, "bad argument");
        mlp_adam_step(h, lr, beta1, beta2, eps, S(stream));
    });
}

// One epoch of training in one call: batch b takes the rows order[b * batch ..
// (b + 1) * batch) of theta_store (FP64, P columns) and of targets (FP32, row
// pitch ldt); each batch runs the gradients of altar_mlp_backward and one
// Adam step with the cosine-decayed rate
// lr_final + (lr - lr_final) (1 + cos(pi it / (total - 1))) / 2, it = it0 + b.
// loss_dev (optional) receives the last batch's loss.
int altar_mlp_train_epoch(altar_mlp *h, const double *theta_store_dev, const float *targets_dev, int ldt,
                          const float *w_dev, const int *order_dev, int nbatches, int batch, int it0, int total,
                          float lr, float lr_final, float beta1, float beta2, float eps, float *loss_dev, void *stream)
{
    return guarded([&] {
        require(h && theta_store_dev && targets_dev && w_dev && order_dev && nbatches > 0 && batch > 0 &&
                    batch <= h->maxB && ldt >= h->nout && total > 0 && lr > 0,
                "bad argument");
        hipStream_t st = S(stream);
        h->xb.reserve((size_t)h->maxB * h->P * 8);
        hip_check(hipMemcpyAsync(h->wk.p, w_dev, h->nout * 4, hipMemcpyDeviceToDevice, st), "weights");
        for (int b = 0; b < nbatches; ++b) {
            const int *sel = order_dev + (size_t)b * batch;
            hip_check(launch_da_pack(theta_store_dev, sel, batch, h->P, h->xb.as<double>(), st), "batch rows");
            hip_check(launch_gather_rows_f32(targets_dev, ldt, sel, batch, h->nout, h->tgt.as<float>(), h->kout, st),
                      "batch targets");
            mlp_staged_gradients(h, h->xb.as<double>(), batch, b == nbatches - 1 ? loss_dev : nullptr, st);
            const double frac = (double)(it0 + b) / (double)std::max(1, total - 1);
            const double cur = lr_final + 0.5 * ((double)lr - lr_final) * (1.0 + std::cos(M_PI * frac));
            mlp_adam_step(h, (float)cur, beta1, beta2, eps, st);
        }
    });
}

// surrogate data llk from predicted normalized coefficients z (C x ldz):
// llk = norm - (cperp + sum_k (e_k + s_k z_k)^2) / 2 (e, s: K values, device FP64)
int altar_surrogate_llk(const float *z_dev, int ldz, int C, int K, const double *e_dev, const double *s_dev,
                        double cperp, double normalization, double *llk_dev, void *stream)
{
    return guarded([&] {
        require(z_dev && e_dev && s_dev && llk_dev && C > 0 && K > 0 && ldz >= K, "bad argument");
        hip_check(launch_surrogate_llk(z_dev, ldz, C, K, e_dev, s_dev, cperp, normalization, llk_dev, S(stream)),
                  "surrogate llk");
    });
}

// ------------------------------------------------------------------ delayed acceptance
int altar_da_screen(const double *cprior_dev, const double *cdata_s_dev, const double *post_s_dev,
                    const int *mask_dev, double beta, int C, uint64_t seed, uint32_t stage, uint32_t step,
                    double *cpost_s_dev, int *index_dev, int *count_dev, unsigned long long *counters_dev,
                    void *stream)
{
    return guarded([&] {
        require(cprior_dev && cdata_s_dev && post_s_dev && mask_dev && cpost_s_dev && index_dev && count_dev &&
                    counters_dev && C > 0,
                "bad argument");
        hip_check(launch_da_screen(cprior_dev, cdata_s_dev, post_s_dev, mask_dev, beta, C, seed, stage, step,
                                   cpost_s_dev, index_dev, count_dev, counters_dev, S(stream)),
                  "da screen");
    });
}

int altar_da_pack(const double *cand_dev, const int *index_dev, int n, int P, double *packed_dev, void *stream)
{
    return guarded([&] {
        require(cand_dev && index_dev && packed_dev && n >= 0 && P > 0, "bad argument");
        hip_check(launch_da_pack(cand_dev, index_dev, n, P, packed_dev, S(stream)), "da pack");
    });
}

int altar_da_accept(double *theta_dev, double *prior_dev, double *data_dev, double *post_dev, double *post_s_dev,
                    const double *cand_dev, const double *cprior_dev, const double *cdata_packed_dev,
                    const double *cpost_s_dev, const int *index_dev, int n, double beta, int P, uint64_t seed,
                    uint32_t stage, uint32_t step, int *decision_dev, unsigned long long *counters_dev,
                    void *stream)
{
    return guarded([&] {
        require(theta_dev && prior_dev && data_dev && post_dev && post_s_dev && cand_dev && cprior_dev &&
                    cdata_packed_dev && cpost_s_dev && index_dev && counters_dev && n >= 0 && P > 0,
                "bad argument");
        hip_check(launch_da_accept(theta_dev, prior_dev, data_dev, post_dev, post_s_dev, cand_dev, cprior_dev,
                                   cdata_packed_dev, cpost_s_dev, index_dev, n, beta, P, seed, stage, step,
                                   decision_dev, counters_dev, S(stream)),
                  "da accept");
    });
}

constexpr uint32_t TRAJ_MAX_SUBSTEPS = 1024;

int altar_da_trajectory_propose(const double *phi_dev, int C, int P, const double *L_dev, uint64_t seed,
                                uint32_t stage, uint32_t step, uint32_t sub, double *xi_dev, double *cand_dev,
                                void *stream)
{
    return guarded([&] {
        require(phi_dev && L_dev && xi_dev && cand_dev && C > 0 && P > 0, "bad argument");
        require(sub < TRAJ_MAX_SUBSTEPS && step < (1u << 22), "sub-step index out of range");
        hip_check(launch_da_traj_propose(phi_dev, C, P, L_dev, seed, stage, step * TRAJ_MAX_SUBSTEPS + sub, xi_dev,
                                         cand_dev, S(stream)),
                  "trajectory propose");
    });
}

int altar_da_trajectory_step(double *phi_dev, double *prior_phi_dev, double *post_phi_dev, const double *cand_dev,
                             const double *cprior_dev, const double *cdata_s_dev, double *cpost_s_dev,
                             const int *mask_dev, double beta, int C, int P, uint64_t seed, uint32_t stage,
                             uint32_t step, uint32_t sub, int *index_dev, unsigned long long *counters_dev,
                             void *stream)
{
    return guarded([&] {
        require(phi_dev && prior_phi_dev && post_phi_dev && cand_dev && cprior_dev && cdata_s_dev && cpost_s_dev &&
                    mask_dev && index_dev && counters_dev && C > 0 && P > 0,
                "bad argument");
        require(sub < TRAJ_MAX_SUBSTEPS && step < (1u << 22), "sub-step index out of range");
        hip_check(launch_da_traj_step(phi_dev, prior_phi_dev, post_phi_dev, cand_dev, cprior_dev, cdata_s_dev,
                                      cpost_s_dev, mask_dev, beta, C, P, seed, stage,
                                      step * TRAJ_MAX_SUBSTEPS + sub, sub == 0 ? 1 : 0, index_dev, counters_dev,
                                      S(stream)),
                  "trajectory step");
    });
}

int altar_da_trajectory_fused(altar_mlp *mlp, const altar_prior *prior, const altar_source *src,
                              const double *theta_dev, const double *prior_theta_dev, double *post_s_dev,
                              double *phi_dev, double *prior_phi_dev, double *post_phi_dev, const double *L_dev,
                              const double *e_dev, const double *s_dev, int K, double cperp, double normalization,
                              double beta, int C, int P, uint64_t seed, uint32_t stage, uint32_t step,
                              int substeps, int *index_dev, unsigned long long *counters_dev, void *stream)
{
    return guarded([&] {
        require(mlp && prior && theta_dev && prior_theta_dev && post_s_dev && phi_dev && prior_phi_dev &&
                    post_phi_dev && L_dev && e_dev && s_dev && index_dev && counters_dev && C > 0 && P > 0,
                "bad argument");
        require(P <= TRAJ_PMAX, "fused trajectories take at most 32 parameters");
        require(P == prior->P && P == mlp->P, "parameter count differs from the prior's or the network's");
        require(mlp->nhid == 256 && mlp->kin == 64 && mlp->kout == 64 && mlp->nlayers >= 2 &&
                    mlp->nlayers <= TRAJ_MAX_LAYERS,
                "fused trajectories need a network of hidden width 256, at most 64 inputs and outputs");
        require(K > 0 && K <= mlp->nout, "K exceeds the network's outputs");
        require(substeps >= 1 && (uint32_t)substeps <= TRAJ_MAX_SUBSTEPS && step < (1u << 22),
                "sub-step count or step index out of range");
        TrajArgs a{};
        a.theta = theta_dev; a.prior_theta = prior_theta_dev; a.post_s = post_s_dev;
        a.phi = phi_dev; a.prior_phi = prior_phi_dev; a.post_phi = post_phi_dev;
        a.L = L_dev;
        a.kind = prior->kind.as<int>(); a.pa = prior->a.as<double>(); a.pb = prior->b.as<double>();
        a.has_cdm = 0;
        if (src && (src->model == ALTAR_CDM || src->model == ALTAR_JOINT)) {
            require(src->have_layout, "layout not set");
            require(src->P == P, "the source's parameter count differs from the sample's");
            a.has_cdm = 1;
            a.cdm = cdm_layout(src);
        }
        for (int l = 0; l < mlp->nlayers; ++l) {
            a.W[l] = mlp->W16[l].as<__bf16>();
            a.bias[l] = mlp->b32[l].as<float>();
        }
        a.nlayers = mlp->nlayers;
        a.lo = mlp->lo.as<double>(); a.hi = mlp->hi.as<double>();
        a.e = e_dev; a.s = s_dev; a.cperp = cperp; a.norm = normalization; a.K = K;
        a.beta = beta; a.C = C; a.P = P; a.seed = seed; a.stage = stage;
        a.step2 = step * TRAJ_MAX_SUBSTEPS;
        a.substeps = substeps;
        a.moved = index_dev + C;
        a.counters = counters_dev;
        hip_check(launch_traj_fused(a, S(stream)), "fused trajectory");
    });
}

int altar_da_trajectory_list(int *index_dev, int C, int *count_dev, void *stream)
{
    return guarded([&] {
        require(index_dev && count_dev && C > 0, "bad argument");
        hip_check(launch_da_traj_list(index_dev, C, count_dev, S(stream)), "trajectory list");
    });
}

}  // extern "C"

extern "C" {

// GPU-resident COV update: median, dbeta grid search and weights on the
// device (out_dev[0] = new beta, out_dev[1] = COV, out_dev[2] = median)
int altar_cov_update_dev(const double *llk_dev, int C, double beta, double tolerance, double target, double *w_dev,
                         double *out_dev, void *stream)
{
    return guarded([&] {
        require(llk_dev && w_dev && out_dev && C > 1, "bad argument");
        require(beta >= 0.0 && beta <= 1.0, "beta outside [0, 1]");
        hip_check(launch_cov_update(llk_dev, C, beta, tolerance, target, w_dev, out_dev, S(stream)), "cov update");
    });
}

// multinomial resampling on the device: src_dev[c] = old row of new row c
// (original-sample order, each repeated by its multiplicity); count_dev (C)
// receives the multiplicities; edges_dev (C + 1 doubles) and offset_dev (C
// ints) are scratch
int altar_cov_resample_dev(const double *w_dev, int C, uint64_t seed, uint32_t stage, double *edges_dev,
                           int *count_dev, int *offset_dev, int *src_dev, void *stream)
{
    return guarded([&] {
        require(w_dev && edges_dev && count_dev && offset_dev && src_dev && C > 0, "bad argument");
        hip_check(launch_cov_resample_src(w_dev, C, seed, stage, edges_dev, count_dev, offset_dev, src_dev, S(stream)),
                  "resample");
    });
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free resampling
// The scratch of altar_cov_resample_dev (the C + 1 bin edges and the C count
// offsets) in a fixed device array, for callers that cannot allocate.  Every
// device the library runs on holds its own copy for the life of the process.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): it passes the address that
// hipGetSymbolAddress returns for the current device.
namespace {

constexpr size_t NOALLOC_EDGE_BYTES = ((size_t)ALTAR_NOALLOC_MAX_CHAINS + 1) * sizeof(double);
constexpr size_t NOALLOC_EDGE_SPAN = (NOALLOC_EDGE_BYTES + 255) / 256 * 256;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_EDGE_SPAN + (size_t)ALTAR_NOALLOC_MAX_CHAINS * sizeof(int);

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the current device's address of scratch (looked up once per device)
void *scratch_address()
{
    int dev = -1;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    require(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device ordinal beyond the supported range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        require(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int altar_cov_resample_dev_noalloc(const double *w_dev, int C, uint64_t seed, uint32_t stage, int *count_dev,
                                   int *src_dev, void *stream)
{
    return guarded([&] {
        require(w_dev && count_dev && src_dev && C > 0, "bad argument");
        require(C <= ALTAR_NOALLOC_MAX_CHAINS,
                "resample (allocation-free): C = " + std::to_string(C) + " exceeds the scratch capacity of " +
                    std::to_string(ALTAR_NOALLOC_MAX_CHAINS) + " chains");
        unsigned char *base = (unsigned char *)scratch_address();
        double *edges = (double *)base;
        int *offset = (int *)(base + NOALLOC_EDGE_SPAN);
        // every scratch word the kernels read is written first in this call
        // (cumsum writes the C + 1 edges, the count scan the C offsets)
        hip_check(launch_cov_resample_src(w_dev, C, seed, stage, edges, count_dev, offset, src_dev, S(stream)),
                  "resample (allocation-free)");
    });
}

}  // extern "C"
