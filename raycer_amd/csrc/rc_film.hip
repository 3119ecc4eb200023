// Film and tone mapping on the device: Film (src/Rendering/Film.cpp:44-69),
// ReinhardToneMapper::apply (src/Rendering/ToneMappers/ReinhardToneMapper.cpp:18-86),
// LinearToneMapper (LinearToneMapper.cpp:12-35), SimpleToneMapper
// (SimpleToneMapper.cpp:12-37) and PassthroughToneMapper (PassthroughToneMapper.cpp:12-22).
//
// The film keeps per pixel the FP32 sums of w*rgb and w plus their Kahan
// compensation terms (32 B + 32 B).  Tone mapping runs in FP64 like the
// reference: for Reinhard a block reduction of sum(log(0.01 + L)) and max(L)
// into per-block partials, a single-block reduction of the partials in a fixed
// order (so repeated runs agree bit for bit), then the per-pixel map with
// MathUtils::fastPow restated on the high word of the double.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>

#include "rc_internal.hpp"

namespace {

constexpr int kBlock = 256;

__global__ void k_film_zero(float4* __restrict__ sum, float4* __restrict__ comp, uint32_t npix) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    sum[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    comp[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// linear = corrected sum / corrected weight (Film::generateToneMappedImage, Film.cpp:63-65)
__device__ __forceinline__ float4 linear_of(float4 s, float4 c) {
    float w = s.w - c.w;
    return make_float4((s.x - c.x) / w, (s.y - c.y) / w, (s.z - c.z) / w, 1.0f);
}

__global__ void k_film_linear(const float4* __restrict__ sum, const float4* __restrict__ comp, uint32_t npix,
                              float4* __restrict__ out) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    out[p] = linear_of(sum[p], comp[p]);
}

__global__ void k_film_sums(const float4* __restrict__ sum, const float4* __restrict__ comp, uint32_t npix,
                            float4* __restrict__ out) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float4 s = sum[p], c = comp[p];
    out[p] = make_float4(s.x - c.x, s.y - c.y, s.z - c.z, s.w - c.w);
}

#pragma clang fp contract(off)

// Color::getLuminance (Color.cpp:130-133) on the float image converted to double.
__device__ __forceinline__ double luminance(float4 c) {
    return 0.2126 * double(c.x) + 0.7152 * double(c.y) + 0.0722 * double(c.z);
}

__global__ __launch_bounds__(kBlock) void k_reduce_partial(const float4* __restrict__ lin, uint32_t npix,
                                                           double2* __restrict__ partial) {
    __shared__ double s_log[kBlock];
    __shared__ double s_max[kBlock];
    double lsum = 0.0, lmax = 0.0;
    for (uint32_t p = blockIdx.x * kBlock + threadIdx.x; p < npix; p += gridDim.x * kBlock) {
        double l = luminance(lin[p]);
        lsum += log(0.01 + l);
        if (l > lmax) lmax = l;
    }
    s_log[threadIdx.x] = lsum;
    s_max[threadIdx.x] = lmax;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
            s_log[threadIdx.x] += s_log[threadIdx.x + o];
            if (s_max[threadIdx.x + o] > s_max[threadIdx.x]) s_max[threadIdx.x] = s_max[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s_log[0], s_max[0]);
}

// MovingAverage::addMeasurement (src/Math/MovingAverage.cpp:28-31) is applied
// to the maximum when averaging is on: avg' = alpha * max + (1 - alpha) * avg.
__global__ __launch_bounds__(kBlock) void k_reduce_final(double2* __restrict__ partial, uint32_t nparts, uint32_t npix,
                                                         uint32_t averaging, double alpha, double average) {
    __shared__ double s_log[kBlock];
    __shared__ double s_max[kBlock];
    double lsum = 0.0, lmax = 1.0;  // maxLuminance starts at 1 (ReinhardToneMapper.cpp:26)
    for (uint32_t i = threadIdx.x; i < nparts; i += kBlock) {
        lsum += partial[i].x;
        if (partial[i].y > lmax) lmax = partial[i].y;
    }
    s_log[threadIdx.x] = lsum;
    s_max[threadIdx.x] = lmax;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) {
            s_log[threadIdx.x] += s_log[threadIdx.x + o];
            if (s_max[threadIdx.x + o] > s_max[threadIdx.x]) s_max[threadIdx.x] = s_max[threadIdx.x + o];
        }
        __syncthreads();
    }
    // partial[nparts] = (log average, max luminance used by the map)
    if (threadIdx.x == 0) {
        double m = s_max[0];
        if (averaging) m = alpha * m + (1.0 - alpha) * average;
        partial[nparts] = make_double2(exp(s_log[0] / double(npix)), m);
    }
}

// MathUtils::fastPow (MathUtils.cpp:56-68): the high 32-bit word is scaled
// about 1072632447 as an integer and the low word cleared.
__device__ __forceinline__ double fast_pow(double a, double b) {
    int hi = __double2hiint(a);
    double x = __dadd_rn(__dmul_rn(b, double(hi - 1072632447)), 1072632447.0);
    return __hiloint2double(int(x), 0);
}

struct MapParams {
    uint32_t type, do_clamp, do_gamma, fast;
    double key, gamma, exposure_scale;
};

// One pixel of the tone map.  Reinhard: scaled luminance s L, mapped
// s L (1 + s L / Lmax^2) / (1 + s L), colour scaled by mapped / original
// luminance.  Linear: c * 2^exposure; simple: c' / (1 + c') of that.  Then
// clamp and gamma as every mapper does.
__global__ __launch_bounds__(kBlock) void k_tone_map(const float4* __restrict__ lin, uint32_t npix,
                                                     const double2* __restrict__ stats, uint32_t nparts, MapParams M,
                                                     float4* __restrict__ out) {
    uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= npix) return;
    float4 c = lin[p];
    if (M.type == RC_TONEMAP_PASSTHROUGH) {
        out[p] = c;
        return;
    }
    double v[3];
    if (M.type == RC_TONEMAP_REINHARD) {
        double2 st = stats[nparts];
        double scale = M.key / st.x;
        double inv2 = 1.0 / (st.y * st.y);
        double l = luminance(c);
        double sl = scale * l;
        double ml = (sl * (1.0 + (sl * inv2))) / (1.0 + sl);
        double cs = ml / l;
        v[0] = double(c.x) * cs;
        v[1] = double(c.y) * cs;
        v[2] = double(c.z) * cs;
    } else {
        v[0] = double(c.x) * M.exposure_scale;
        v[1] = double(c.y) * M.exposure_scale;
        v[2] = double(c.z) * M.exposure_scale;
        if (M.type == RC_TONEMAP_SIMPLE)
            for (int k = 0; k < 3; ++k) v[k] = v[k] / (1.0 + v[k]);
    }
    double ig = 1.0 / M.gamma;
    for (int k = 0; k < 3; ++k) {
        // Color::clamped: std::max(0, std::min(c, 1)), which maps NaN to 0
        if (M.do_clamp) {
            double m = (1.0 < v[k]) ? 1.0 : v[k];
            v[k] = (0.0 < m) ? m : 0.0;
        }
        if (M.do_gamma) v[k] = M.fast ? fast_pow(v[k], ig) : pow(v[k], ig);
    }
    out[p] = make_float4(float(v[0]), float(v[1]), float(v[2]), 1.0f);
}

#pragma clang fp contract(on)

// MathUtils::fastPow on the host, for 2^exposure.
double host_fast_pow(double a, double b) {
    uint64_t bits;
    std::memcpy(&bits, &a, sizeof bits);
    const int hi = int(int32_t(bits >> 32));
    const double m = b * double(hi - 1072632447);  // two statements: no fused multiply-add
    const int x = int(m + 1072632447.0);
    bits = uint64_t(uint32_t(x)) << 32;
    double r;
    std::memcpy(&r, &bits, sizeof r);
    return r;
}

int check_film(rc_ctx* c, const rc_film* f, const char* fn) {
    RC_TRY(rc_check_ctx(c, fn));
    if (!f || f->ctx != c) return rc_host_fail(RC_ERR_INVALID, std::string(fn) + ": bad film");
    return RC_OK;
}

}  // namespace

extern "C" int rc_film_create(rc_ctx* c, uint32_t w, uint32_t h, rc_film** out) {
    RC_TRY(rc_check_ctx(c, "rc_film_create"));
    if (!out || w < 2 || h < 2 || size_t(w) * h > 0x7fffffffull)
        return rc_host_fail(RC_ERR_INVALID, "rc_film_create: bad size");
    RC_TRY_HIP(hipSetDevice(c->device));
    rc_film* f = new (std::nothrow) rc_film();
    if (!f) return rc_host_fail(RC_ERR_ALLOC, "rc_film_create: out of memory");
    f->ctx = c;
    f->width = w;
    f->height = h;
    size_t npix = size_t(w) * h;
    if (hipMalloc(&f->sum, npix * sizeof(float4)) != hipSuccess || hipMalloc(&f->comp, npix * sizeof(float4)) != hipSuccess) {
        if (f->sum) (void)hipFree(f->sum);
        delete f;
        return rc_host_fail(RC_ERR_ALLOC, "rc_film_create: device allocation failed");
    }
    int rc = rc_film_clear(c, f);
    if (rc != RC_OK) {
        rc_film_free(f);
        return rc;
    }
    *out = f;
    return RC_OK;
}

extern "C" int rc_film_free(rc_film* f) {
    if (!f) return RC_OK;
    (void)hipStreamSynchronize(f->ctx->stream);
    (void)hipFree(f->sum);
    (void)hipFree(f->comp);
    f->linear.release();
    f->partial.release();
    delete f;
    return RC_OK;
}

// Film::clear (Film.cpp:44-47)
extern "C" int rc_film_clear(rc_ctx* c, rc_film* f) {
    RC_TRY(check_film(c, f, "rc_film_clear"));
    uint32_t npix = f->width * f->height;
    k_film_zero<<<(npix + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(f->sum, f->comp, npix);
    RC_TRY_HIP(hipGetLastError());
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

extern "C" int rc_film_download_linear(rc_ctx* c, const rc_film* f, float* rgba) {
    RC_TRY(check_film(c, f, "rc_film_download_linear"));
    if (!rgba) return rc_host_fail(RC_ERR_INVALID, "rc_film_download_linear: null output");
    uint32_t npix = f->width * f->height;
    rc_film* mf = const_cast<rc_film*>(f);
    RC_TRY(mf->linear.ensure(size_t(npix) * sizeof(float4)));
    k_film_linear<<<(npix + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(f->sum, f->comp, npix, mf->linear.as<float4>());
    RC_TRY_HIP(hipGetLastError());
    RC_TRY_HIP(hipMemcpyAsync(rgba, mf->linear.p, size_t(npix) * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

extern "C" int rc_film_download_sums(rc_ctx* c, const rc_film* f, float* sums) {
    RC_TRY(check_film(c, f, "rc_film_download_sums"));
    if (!sums) return rc_host_fail(RC_ERR_INVALID, "rc_film_download_sums: null output");
    uint32_t npix = f->width * f->height;
    rc_film* mf = const_cast<rc_film*>(f);
    RC_TRY(mf->linear.ensure(size_t(npix) * sizeof(float4)));
    k_film_sums<<<(npix + kBlock - 1) / kBlock, kBlock, 0, c->stream>>>(f->sum, f->comp, npix, mf->linear.as<float4>());
    RC_TRY_HIP(hipGetLastError());
    RC_TRY_HIP(hipMemcpyAsync(sums, mf->linear.p, size_t(npix) * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    return RC_OK;
}

extern "C" int rc_film_tonemap(rc_ctx* c, const rc_film* f, const rc_tonemap* tm, float* rgba, double* log_avg,
                               double* max_lum) {
    RC_TRY(check_film(c, f, "rc_film_tonemap"));
    if (!tm || !rgba) return rc_host_fail(RC_ERR_INVALID, "rc_film_tonemap: null argument");
    if (tm->type > RC_TONEMAP_PASSTHROUGH) return rc_host_fail(RC_ERR_INVALID, "rc_film_tonemap: unknown tone mapper");
    const bool reinhard = tm->type == RC_TONEMAP_REINHARD;
    if (!(tm->gamma > 0.0f) || (reinhard && !(tm->key > 0.0f)))
        return rc_host_fail(RC_ERR_INVALID, "rc_film_tonemap: bad key/gamma");
    if (reinhard && tm->enable_averaging && !(tm->averaging_alpha > 0.0f && tm->averaging_alpha <= 1.0f))
        return rc_host_fail(RC_ERR_INVALID, "rc_film_tonemap: averaging_alpha must be in (0, 1]");
    if (!std::isfinite(tm->exposure)) return rc_host_fail(RC_ERR_INVALID, "rc_film_tonemap: bad exposure");
    uint32_t npix = f->width * f->height;
    rc_film* mf = const_cast<rc_film*>(f);
    const uint32_t nparts = std::min<uint32_t>(1024, (npix + kBlock - 1) / kBlock);
    RC_TRY(mf->linear.ensure(size_t(npix) * sizeof(float4) * 2));
    RC_TRY(mf->partial.ensure(size_t(nparts + 1) * sizeof(double2)));
    float4* lin = mf->linear.as<float4>();
    float4* outd = lin + npix;
    double2* part = mf->partial.as<double2>();
    uint32_t grid = (npix + kBlock - 1) / kBlock;
    const bool averaging = reinhard && tm->enable_averaging;
    MapParams M;
    M.type = tm->type;
    M.do_clamp = tm->clamp;
    M.do_gamma = tm->apply_gamma;
    M.fast = tm->fast_pow;
    M.key = double(tm->key);
    M.gamma = double(tm->gamma);
    M.exposure_scale = tm->fast_pow ? host_fast_pow(2.0, double(tm->exposure)) : std::exp2(double(tm->exposure));
    k_film_linear<<<grid, kBlock, 0, c->stream>>>(f->sum, f->comp, npix, lin);
    if (reinhard) {
        k_reduce_partial<<<nparts, kBlock, 0, c->stream>>>(lin, npix, part);
        k_reduce_final<<<1, kBlock, 0, c->stream>>>(part, nparts, npix, averaging ? 1u : 0u,
                                                    double(tm->averaging_alpha), f->max_lum_average);
    }
    k_tone_map<<<grid, kBlock, 0, c->stream>>>(lin, npix, part, nparts, M, outd);
    RC_TRY_HIP(hipGetLastError());
    RC_TRY_HIP(hipMemcpyAsync(rgba, outd, size_t(npix) * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    double2 st = make_double2(0.0, 0.0);
    if (reinhard) RC_TRY_HIP(hipMemcpyAsync(&st, part + nparts, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    RC_TRY_HIP(hipStreamSynchronize(c->stream));
    if (averaging) mf->max_lum_average = st.y;
    if (log_avg) *log_avg = st.x;
    if (max_lum) *max_lum = st.y;
    return RC_OK;
}
