// libsfk C ABI: context, storage, camera, sort/scan primitives and the
// render / seg orchestration (the host side of src/render.rs, src/seg/mod.rs,
// src/seg/beta.rs and crates/sort, over one HIP stream).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>

#include "sfk_internal.h"

using namespace sfk;

namespace sfk {

int fail(sfk_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int hip_fail(sfk_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, e == hipErrorOutOfMemory ? SFK_E_NOMEM : SFK_E_HIP, "%s: %s (%d)", what,
                hipGetErrorString(e), (int)e);
}

hipEvent_t prof_event(sfk_ctx* c) {
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

}  // namespace sfk

extern "C" int sfk_prof_enable(sfk_ctx* c, int on) {
    if (!c) return SFK_E_ARG;
    c->prof = on != 0;
    return SFK_OK;
}

// Aggregate the recorded launches by name: "name count total_ms\n" lines.
extern "C" int sfk_prof_report(sfk_ctx* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return SFK_E_ARG;
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<uint32_t> cnt;
    for (auto& r : c->prof_recs) {
        float t = 0.0f;
        hipEventElapsedTime(&t, r.a, r.b);
        size_t k = 0;
        while (k < names.size() && names[k] != r.name) ++k;
        if (k == names.size()) {
            names.emplace_back(r.name);
            ms.push_back(0.0);
            cnt.push_back(0);
        }
        ms[k] += t;
        cnt[k] += 1;
        c->prof_pool.push_back(r.a);
        c->prof_pool.push_back(r.b);
    }
    c->prof_recs.clear();
    std::string out;
    for (size_t k = 0; k < names.size(); ++k) {
        char line[256];
        snprintf(line, sizeof line, "%s %u %.6f\n", names[k].c_str(), cnt[k], ms[k]);
        out += line;
    }
    snprintf(buf, (size_t)len, "%s", out.c_str());
    return (int)out.size() < len ? SFK_OK : SFK_E_CAPACITY;
}

// ----------------------------------------------------------------- context

// Fixed reduction workspace of sfk_eig_device_noalloc: one f64 partial per
// EIG block of the largest admitted splat count.  Each device holds its own
// copy (512 KiB) for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress gives for that device.
static __device__ __align__(256) double scratch[SFK_EIG_NOALLOC_MAX_SPLATS / 1024];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_fixed_mu;
double* g_fixed_ptr[kMaxDevices] = {};

// This device's address of `scratch`, looked up once per device with the
// device current (the caller has set it).
hipError_t fixed_scratch(int device, double** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_fixed_mu);
    if (!g_fixed_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_fixed_ptr[device] = static_cast<double*>(p);
    }
    *out = g_fixed_ptr[device];
    return hipSuccess;
}
}  // namespace

extern "C" int sfk_ctx_create(int device, sfk_ctx** out) {
    if (!out) return SFK_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return SFK_E_HIP;
    if (hipSetDevice(device) != hipSuccess) return SFK_E_HIP;
    auto* c = new (std::nothrow) sfk_ctx();
    if (!c) return SFK_E_NOMEM;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipMalloc(&c->eig_out, sizeof(double)) != hipSuccess) {
        delete c;
        return SFK_E_HIP;
    }
    // A failed lookup only disables sfk_eig_device_noalloc, which reports it.
    c->eig_fixed_err = fixed_scratch(device, &c->eig_fixed);
    *out = c;
    return SFK_OK;
}

extern "C" int sfk_eig_device_noalloc(sfk_ctx* c, const float* a, const float* b,
                                      const uint32_t* resp, uint32_t n, float scale,
                                      double* d_out) {
    if (!c || !d_out || (n && (!a || !b || !resp))) return SFK_E_ARG;
    if (n > SFK_EIG_NOALLOC_MAX_SPLATS)
        return fail(c, SFK_E_CAPACITY, "eig (no alloc): %u splats, at most %u", n,
                    (uint32_t)SFK_EIG_NOALLOC_MAX_SPLATS);
    constexpr uint32_t kFixedPartials = sizeof(scratch) / sizeof(double);
    if (eig_num_partials(n) > kFixedPartials)
        return fail(c, SFK_E_CAPACITY, "eig (no alloc): %u partials, room for %u",
                    eig_num_partials(n), kFixedPartials);
    if (!c->eig_fixed) return hip_fail(c, c->eig_fixed_err, "eig (no alloc): workspace address");
    // Every partial the final sum reads is written by this call's first
    // kernel, so nothing needs clearing between calls.
    SFK_HIP(c, launch_eig(c->stream, a, b, resp, n, scale, c->eig_fixed, d_out));
    return SFK_OK;
}

extern "C" void sfk_ctx_destroy(sfk_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    if (c->eig_partials) hipFree(c->eig_partials);
    if (c->eig_out) hipFree(c->eig_out);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* sfk_last_error(const sfk_ctx* c) { return c ? c->err.c_str() : ""; }
extern "C" void* sfk_ctx_stream(sfk_ctx* c) { return c ? (void*)c->stream : nullptr; }

extern "C" int sfk_sync(sfk_ctx* c) {
    if (!c) return SFK_E_ARG;
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    return SFK_OK;
}

extern "C" int sfk_device_name(sfk_ctx* c, char* buf, int len) {
    if (!c || !buf || len <= 0) return SFK_E_ARG;
    hipDeviceProp_t p;
    SFK_HIP(c, hipGetDeviceProperties(&p, c->device));
    snprintf(buf, (size_t)len, "%s %s cu=%d", p.name, p.gcnArchName, p.multiProcessorCount);
    return SFK_OK;
}

extern "C" int sfk_alloc(sfk_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return SFK_E_ARG;
    *out = nullptr;
    SFK_HIP(c, hipMalloc(out, bytes ? bytes : 4));
    return SFK_OK;
}

extern "C" int sfk_free(sfk_ctx* c, void* p) {
    if (!c) return SFK_E_ARG;
    if (p) SFK_HIP(c, hipFree(p));
    return SFK_OK;
}

extern "C" int sfk_upload(sfk_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return SFK_E_ARG;
    if (!bytes) return SFK_OK;
    SFK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    return SFK_OK;
}

extern "C" int sfk_download(sfk_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return SFK_E_ARG;
    if (!bytes) return SFK_OK;
    SFK_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    SFK_HIP(c, hipStreamSynchronize(c->stream));
    return SFK_OK;
}

extern "C" int sfk_fill_u32(sfk_ctx* c, uint32_t* buf, uint32_t v, size_t count) {
    if (!c || (count && !buf)) return SFK_E_ARG;
    SFK_HIP(c, launch_fill_u32(c->stream, buf, v, count));
    return SFK_OK;
}

extern "C" int sfk_copy_u32(sfk_ctx* c, uint32_t* dst, const uint32_t* src, size_t count) {
    if (!c || (count && (!dst || !src)) || count % 4) return SFK_E_ARG;
    SFK_PROF(c, "copy_u32");
    SFK_HIP(c, launch_copy_u32(c->stream, dst, src, count));
    return SFK_OK;
}

extern "C" int sfk_timer_start(sfk_ctx* c) {
    if (!c) return SFK_E_ARG;
    SFK_HIP(c, hipEventRecord(c->ev0, c->stream));
    return SFK_OK;
}

extern "C" int sfk_timer_stop_ms(sfk_ctx* c, float* ms) {
    if (!c || !ms) return SFK_E_ARG;
    SFK_HIP(c, hipEventRecord(c->ev1, c->stream));
    SFK_HIP(c, hipEventSynchronize(c->ev1));
    SFK_HIP(c, hipEventElapsedTime(ms, c->ev0, c->ev1));
    return SFK_OK;
}

// ------------------------------------------------------------------ camera

// Rotation matrix columns of a unit quaternion (glam's Mat3::from_quat form).
static void quat_cols(const float q[4], float c0[3], float c1[3], float c2[3]) {
    float x = q[0], y = q[1], z = q[2], w = q[3];
    float x2 = x + x, y2 = y + y, z2 = z + z;
    float xx = x * x2, xy = x * y2, xz = x * z2;
    float yy = y * y2, yz = y * z2, zz = z * z2;
    float wx = w * x2, wy = w * y2, wz = w * z2;
    c0[0] = 1.0f - (yy + zz); c0[1] = xy + wz;          c0[2] = xz - wy;
    c1[0] = xy - wz;          c1[1] = 1.0f - (xx + zz); c1[2] = yz + wx;
    c2[0] = xz + wy;          c2[1] = yz - wx;          c2[2] = 1.0f - (xx + yy);
}

static float hdot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// CameraViewLaunch::for_camera (src/project.rs:76-96); w2c = (R^T, -R^T p),
// the exact rigid inverse of camera.rs:42-44.
extern "C" int sfk_camera_view(const sfk_camera* cam, uint32_t w, uint32_t h, uint32_t tile_w,
                               sfk_view* out) {
    if (!cam || !out || w == 0 || h == 0 || !(tile_w == 8 || tile_w == 16 || tile_w == 32))
        return SFK_E_ARG;
    quat_cols(cam->rotation, out->rot0, out->rot1, out->rot2);
    out->trans[0] = -hdot3(out->rot0, cam->position);
    out->trans[1] = -hdot3(out->rot1, cam->position);
    out->trans[2] = -hdot3(out->rot2, cam->position);
    out->focal[0] = ((float)w * 0.5f) / tanf(cam->fov[0] * 0.5f);
    out->focal[1] = ((float)h * 0.5f) / tanf(cam->fov[1] * 0.5f);
    for (int k = 0; k < 3; ++k) out->world_pos[k] = cam->position[k];
    out->img[0] = (float)w;
    out->img[1] = (float)h;
    out->tile_bounds[0] = (float)((w + tile_w - 1) / tile_w);
    out->tile_bounds[1] = (float)((h + tile_w - 1) / tile_w);
    return SFK_OK;
}

// Camera::look_at (src/camera.rs:57-63): minimal arc from +Z; antiparallel
// turns pi about +Y.
extern "C" int sfk_look_at(const float pos[3], const float target[3], float q[4]) {
    if (!pos || !target || !q) return SFK_E_ARG;
    float d[3] = {target[0] - pos[0], target[1] - pos[1], target[2] - pos[2]};
    float inv = 1.0f / sqrtf(hdot3(d, d));
    float dir[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
    const float eps = 1.1920929e-7f;
    float dp = dir[2];
    if (dp > 1.0f - 2.0f * eps) {
        q[0] = 0.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 1.0f;
    } else if (dp < -1.0f + 2.0f * eps) {
        float half = 3.14159265358979323846f * 0.5f;
        q[0] = 0.0f; q[1] = sinf(half); q[2] = 0.0f; q[3] = cosf(half);
    } else {
        float c[4] = {-dir[1], dir[0], 0.0f, 1.0f + dp};
        float len = sqrtf((c[0] * c[0] + c[1] * c[1]) + (c[2] * c[2] + c[3] * c[3]));
        float r = 1.0f / len;
        q[0] = c[0] * r; q[1] = c[1] * r; q[2] = c[2] * r; q[3] = c[3] * r;
    }
    return SFK_OK;
}

// --------------------------------------------------------------- sort/scan

extern "C" uint32_t sfk_bits_for(uint32_t m) {
    uint32_t v = m == 0 ? 0 : m - 1, b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

struct sfk_radix_scratch {
    sfk_ctx* ctx = nullptr;
    uint32_t max_elems = 0, key_bytes = 4, val_bytes = 4, digit_bits = 8, max_blocks = 0;
    void* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint32_t* hist = nullptr;
    uint32_t* totals = nullptr;
    // u64 bucket finish (launch_bucket_split / launch_bucket_sort): on unless
    // switched off; buffers allocated for u64 scratches only.
    bool bucket_finish = true;
    uint32_t* bucket_first = nullptr;  // 65536
    uint32_t* bucket_last = nullptr;   // 65536
    uint32_t* bucket_size = nullptr;   // 65536, then bucket_size_groups() group maxima
    uint32_t* bucket_max_host = nullptr;  // pinned, bucket_size_groups()
};

static void radix_release(sfk_radix_scratch* s) {
    if (!s) return;
    for (int i = 0; i < 2; ++i) {
        if (s->keys[i]) hipFree(s->keys[i]);
        if (s->vals[i]) hipFree(s->vals[i]);
        s->keys[i] = nullptr;
        s->vals[i] = nullptr;
    }
    if (s->hist) hipFree(s->hist);
    if (s->totals) hipFree(s->totals);
    if (s->bucket_first) hipFree(s->bucket_first);
    if (s->bucket_last) hipFree(s->bucket_last);
    if (s->bucket_size) hipFree(s->bucket_size);
    if (s->bucket_max_host) hipHostFree(s->bucket_max_host);
    s->hist = s->totals = nullptr;
    s->bucket_first = s->bucket_last = s->bucket_size = s->bucket_max_host = nullptr;
}

static int radix_alloc(sfk_ctx* c, sfk_radix_scratch* s, uint32_t max_elems) {
    size_t m = max_elems ? max_elems : 1;
    s->max_elems = max_elems;
    s->max_blocks = radix_max_blocks(max_elems);
    for (int i = 0; i < 2; ++i) {
        SFK_HIP(c, hipMalloc(&s->keys[i], m * s->key_bytes));
        SFK_HIP(c, hipMalloc(&s->vals[i], m * s->val_bytes));
    }
    // Room for up to 256 segments (each may round its last block up).
    SFK_HIP(c, hipMalloc(&s->hist, ((size_t)s->max_blocks + 256) * 256 * 4));
    SFK_HIP(c, hipMalloc(&s->totals, 256 * 256 * 4));
    if (s->key_bytes == 8 && max_elems >= kBucketMinKeys) {
        SFK_HIP(c, hipMalloc(&s->bucket_first, 65536 * 4));
        SFK_HIP(c, hipMalloc(&s->bucket_last, 65536 * 4));
        SFK_HIP(c, hipMalloc(&s->bucket_size, (65536 + (size_t)bucket_size_groups()) * 4));
        SFK_HIP(c, hipHostMalloc(&s->bucket_max_host, (size_t)bucket_size_groups() * 4,
                                 hipHostMallocDefault));
    }
    return SFK_OK;
}

int sfk::radix_scratch_create_ex(sfk_ctx* c, uint32_t max_elems, uint32_t key_bytes,
                                 uint32_t val_bytes, sfk_radix_scratch** out) {
    if (!c || !out || (key_bytes != 4 && key_bytes != 8) || (val_bytes != 4 && val_bytes != 8) ||
        (val_bytes == 8 && key_bytes != 4))
        return fail(c, SFK_E_ARG, "radix scratch: bad args");
    *out = nullptr;
    auto* s = new (std::nothrow) sfk_radix_scratch();
    if (!s) return SFK_E_NOMEM;
    s->ctx = c;
    s->key_bytes = key_bytes;
    s->val_bytes = val_bytes;
    int rc = radix_alloc(c, s, max_elems);
    if (rc) {
        radix_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return SFK_OK;
}

extern "C" int sfk_radix_scratch_create(sfk_ctx* c, uint32_t max_elems, uint32_t key_bytes,
                                        sfk_radix_scratch** out) {
    return sfk::radix_scratch_create_ex(c, max_elems, key_bytes, 4, out);
}

extern "C" void sfk_radix_scratch_destroy(sfk_radix_scratch* s) {
    if (!s) return;
    radix_release(s);
    delete s;
}

// An argsort reads its input in the first pass only, and that pass writes
// keys[0] / vals[0]: the second buffer pair may hold the input itself.
const uint32_t* sfk::radix_scratch_totals(sfk_radix_scratch* s, uint32_t* digit_bits) {
    if (digit_bits) *digit_bits = s ? s->digit_bits : 0u;
    return s ? s->totals : nullptr;
}

int sfk::radix_scratch_spare_u32x2(sfk_radix_scratch* s, uint32_t** keys, uint2** vals) {
    if (!s || s->key_bytes != 4 || s->val_bytes != 8) return SFK_E_ARG;
    *keys = static_cast<uint32_t*>(s->keys[1]);
    *vals = reinterpret_cast<uint2*>(s->vals[1]);
    return SFK_OK;
}


extern "C" int sfk_radix_set_bucket_finish(sfk_radix_scratch* s, int on) {
    if (!s) return SFK_E_ARG;
    s->bucket_finish = on != 0;
    return SFK_OK;
}

extern "C" int sfk_radix_set_digit_bits(sfk_radix_scratch* s, uint32_t bits) {
    if (!s || bits < 4 || bits > 8) return SFK_E_ARG;
    s->digit_bits = bits;
    return SFK_OK;
}

template <typename K, typename V = uint32_t>
static int argsort_impl(sfk_ctx* c, sfk_radix_scratch* s, const K* keys, const V* vals,
                        uint32_t n, const uint32_t* d_n, uint32_t bits, int write_keys,
                        K** out_keys, V** out_vals, uint32_t nseg = 1,
                        const uint32_t* payload = nullptr, const KeyRebase* rebase = nullptr) {
    if (!c || !s || !out_vals || (n && !keys)) return fail(c, SFK_E_ARG, "argsort: bad args");
    if (sizeof(K) != s->key_bytes) return fail(c, SFK_E_ARG, "argsort: scratch key width %u", s->key_bytes);
    if (sizeof(V) > s->val_bytes) return fail(c, SFK_E_ARG, "argsort: scratch value width %u", s->val_bytes);
    if (bits > 8 * sizeof(K)) return fail(c, SFK_E_ARG, "argsort: %u bits on %zu-byte keys", bits, sizeof(K));
    if (nseg == 0 || nseg > 256 || (nseg > 1 && d_n))
        return fail(c, SFK_E_ARG, "argsort: %u segments", nseg);
    const uint64_t total = (uint64_t)n * nseg;  // n = keys per segment
    if (total > s->max_elems)
        return fail(c, SFK_E_CAPACITY, "radix scratch undersized: sorted %llu elements, capacity %u",
                    (unsigned long long)total, s->max_elems);
    // n <= 1 or bits == 0: nothing moves (sort.rs:319-321).
    if ((n <= 1 || bits == 0) && vals) {
        if (out_keys) *out_keys = const_cast<K*>(keys);
        *out_vals = const_cast<V*>(vals);
        return SFK_OK;
    }
    if (n == 0) {
        if (out_keys) *out_keys = const_cast<K*>(keys);
        *out_vals = static_cast<V*>(static_cast<void*>(s->vals[0]));
        return SFK_OK;
    }
    // Without values the identity permutation is wanted: a zero-width pass
    // is a stable copy that materializes it.
    if (n <= 1 || bits == 0) bits = 0;
    if (sizeof(K) == 8 && s->bucket_finish && s->bucket_first && !d_n && nseg == 1 &&
        bits >= 24 && n >= kBucketMinKeys) {
        // Two stable passes on the top 16 bits group the keys by them; if
        // every group fits on chip, one launch sorts each by the rest.
        SortArgs a;
        memset(&a, 0, sizeof a);
        a.n_max = n;
        a.nseg = 1;
        a.hist = s->hist;
        a.kpb = radix_keys_per_block(n);
        a.stride = s->max_blocks;
        a.totals = s->totals;
        a.width = 8;
        for (uint32_t p = 0; p < 2; ++p) {
            a.keys_in = p == 0 ? static_cast<const void*>(keys) : s->keys[0];
            a.vals_in = p == 0 ? static_cast<const void*>(vals) : s->vals[0];
            a.keys_out = s->keys[p];
            a.vals_out = s->vals[p];
            a.shift = bits - 16 + 8 * p;
            SFK_PROF(c, "radix_pass_u64");
            hipError_t e = launch_radix_pass(c->stream, a, 8, sizeof(V), 8);
            if (e != hipSuccess) return hip_fail(c, e, "radix pass");
        }
        const uint64_t* grouped = static_cast<const uint64_t*>(s->keys[1]);
        const uint32_t ng = bucket_size_groups();
        {
            SFK_PROF(c, "radix_bucket_split");
            hipError_t e = launch_bucket_split(c->stream, grouped, n, bits - 16, s->bucket_first,
                                               s->bucket_last, s->bucket_size,
                                               s->bucket_size + 65536);
            if (e != hipSuccess) return hip_fail(c, e, "radix bucket split");
        }
        {
            // Queued before the size check is read back: in the usual case
            // the host waits once, for the finished sort.
            SFK_PROF(c, "radix_bucket_sort");
            hipError_t e = launch_bucket_sort(
                c->stream, grouped, s->vals[1],
                write_keys ? static_cast<uint64_t*>(s->keys[0]) : nullptr, s->vals[0],
                s->bucket_first, s->bucket_size, bits - 16);
            if (e == hipSuccess)
                e = hipMemcpyAsync(s->bucket_max_host, s->bucket_size + 65536, (size_t)ng * 4,
                                   hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "radix bucket sort");
        }
        uint32_t mx = 0;
        for (uint32_t g = 0; g < ng; ++g) mx = std::max(mx, s->bucket_max_host[g]);
        if (mx <= bucket_cap()) {
            if (out_keys) *out_keys = static_cast<K*>(s->keys[0]);
            *out_vals = reinterpret_cast<V*>(s->vals[0]);
            return SFK_OK;
        }
        // A bucket too large for one wavefront (the bucket sort left it
        // alone): the plain passes below, from the input again.
    }
    const K* cur_k = keys;
    const V* cur_v = vals;
    int dst = 0;
    // As few passes as digits of at most digit_bits allow, with the bits
    // spread evenly over them (13 bits: 7 + 6, not 8 + 5): fewer bins per
    // pass mean longer coalesced runs in the scatter.
    const uint32_t db = s->digit_bits;
    const uint32_t passes = bits == 0 ? 1u : (bits + db - 1) / db;
    const uint32_t wlo = bits / passes, wextra = bits % passes;
    SortArgs a;
    a.n_max = nseg > 1 ? 0u : n;
    a.d_n = d_n;
    a.nseg = nseg;
    a.seg_len = nseg > 1 ? n : 0u;
    a.hist = s->hist;
    a.kpb = radix_keys_per_block(total);
    a.stride = nseg > 1 ? div_up(n, a.kpb) : s->max_blocks;
    a.totals = s->totals;
    a.rebase_mode = 0;
    a.rebase_base = 0;
    a.rebase_bases = nullptr;
    uint32_t shift = 0;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t width = wlo + (p < wextra ? 1u : 0u);
        bool last = p + 1 == passes;
        a.keys_in = cur_k;
        a.vals_in = cur_v;
        a.keys_out = (!last || write_keys) ? s->keys[dst] : nullptr;
        a.vals_out = s->vals[dst];
        a.payload = payload;
        a.shift = shift;
        a.width = width;
        a.rebase_mode = (p == 0 && rebase && sizeof(K) == 4) ? (rebase->bases ? 2u : 1u) : 0u;
        if (a.rebase_mode) {
            a.rebase_base = rebase->base;
            a.rebase_bases = rebase->bases;
        }
        SFK_PROF(c, sizeof(K) == 4 ? "radix_pass_u32" : "radix_pass_u64");
        hipError_t e = launch_radix_pass(c->stream, a, sizeof(K), sizeof(V),
                                         wlo + (wextra ? 1u : 0u));
        if (e != hipSuccess) return hip_fail(c, e, "radix pass");
        shift += width;
        cur_k = static_cast<K*>(s->keys[dst]);
        cur_v = static_cast<const V*>(a.vals_out);
        dst ^= 1;
    }
    if (out_keys) *out_keys = const_cast<K*>(cur_k);
    *out_vals = const_cast<V*>(cur_v);
    return SFK_OK;
}

int sfk::argsort_payload_u32(sfk_ctx* c, sfk_radix_scratch* s, const uint32_t* keys,
                             const uint32_t* payload, uint32_t n, uint32_t nseg, uint32_t bits,
                             uint2** out, const KeyRebase* rebase) {
    if (!payload) return fail(c, SFK_E_ARG, "argsort: no payload");
    return argsort_impl<uint32_t, uint2>(c, s, keys, nullptr, n, nullptr, bits, 0, nullptr, out,
                                         nseg, payload, rebase);
}

static uint32_t f32_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

// project_geom computes cz = dot(rot2, mean) + trans.z in f32 and culls
// cz <= 0.1.  Over the AABB the exact value lies in [lo, hi]; the f32 one
// within a few ulps of the magnitudes involved, covered by a relative
// margin of 1e-5.
bool sfk::depth_rebase(const sfk_view& v, const float bmin[3], const float bmax[3],
                       uint32_t* base) {
    double lo = v.trans[2], hi = v.trans[2], mag = fabs((double)v.trans[2]);
    for (int i = 0; i < 3; ++i) {
        const double r = v.rot2[i], p = r * bmin[i], q = r * bmax[i];
        lo += std::min(p, q);
        hi += std::max(p, q);
        mag += fabs(r) * std::max(fabs((double)bmin[i]), fabs((double)bmax[i]));
    }
    const double err = 1e-5 * mag + 1e-30;
    const float near = 0.1f;
    float flo = (float)(lo - err);
    if ((double)flo > lo - err) flo = nextafterf(flo, -INFINITY);
    float fhi = (float)(hi + err);
    if ((double)fhi < hi + err) fhi = nextafterf(fhi, INFINITY);
    const uint32_t kmin = flo > near ? f32_bits(flo) : f32_bits(near);
    const uint32_t kmax = fhi > near ? f32_bits(fhi) : kmin;
    *base = kmin;
    return kmax - kmin < 0xfffffeu && std::isfinite(fhi);
}

int sfk::argsort_u32_vals8(sfk_ctx* c, sfk_radix_scratch* s, const uint32_t* keys,
                           const uint2* vals, uint32_t n, const uint32_t* d_n, uint32_t bits,
                           uint32_t** out_keys, uint2** out_vals) {
    if (!vals) return fail(c, SFK_E_ARG, "argsort: no values");
    return argsort_impl<uint32_t, uint2>(c, s, keys, vals, n, d_n, bits, 1, out_keys, out_vals);
}

extern "C" int sfk_argsort_u32(sfk_ctx* c, sfk_radix_scratch* s, const uint32_t* keys,
                               const uint32_t* vals, uint32_t n, const uint32_t* d_n,
                               uint32_t bits, int write_keys, uint32_t** out_keys,
                               uint32_t** out_vals) {
    return argsort_impl<uint32_t>(c, s, keys, vals, n, d_n, bits, write_keys, out_keys, out_vals);
}

extern "C" int sfk_argsort_u64(sfk_ctx* c, sfk_radix_scratch* s, const uint64_t* keys,
                               const uint32_t* vals, uint32_t n, const uint32_t* d_n,
                               uint32_t bits, int write_keys, uint64_t** out_keys,
                               uint32_t** out_vals) {
    return argsort_impl<uint64_t>(c, s, keys, vals, n, d_n, bits, write_keys, out_keys, out_vals);
}

struct sfk_scan_scratch {
    uint32_t max_elems = 0;
    uint32_t* sums = nullptr;
};

extern "C" int sfk_scan_scratch_create(sfk_ctx* c, uint32_t max_elems, sfk_scan_scratch** out) {
    if (!c || !out) return SFK_E_ARG;
    *out = nullptr;
    auto* s = new (std::nothrow) sfk_scan_scratch();
    if (!s) return SFK_E_NOMEM;
    s->max_elems = max_elems;
    hipError_t e = hipMalloc(&s->sums, (size_t)div_up(max_elems ? max_elems : 1, kScanBlock) * 4);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(c, e, "scan scratch");
    }
    *out = s;
    return SFK_OK;
}

uint32_t* sfk::scan_scratch_sums(sfk_scan_scratch* s) { return s ? s->sums : nullptr; }

extern "C" void sfk_scan_scratch_destroy(sfk_scan_scratch* s) {
    if (!s) return;
    if (s->sums) hipFree(s->sums);
    delete s;
}

extern "C" int sfk_exclusive_scan_gather(sfk_ctx* c, const uint32_t* gids, const uint32_t* counts,
                                         uint32_t* offsets, uint32_t n, sfk_scan_scratch* s) {
    if (!c || !s) return SFK_E_ARG;
    if (n == 0) return SFK_OK;  // scan.rs:226-228
    if (!gids || !counts || !offsets) return fail(c, SFK_E_ARG, "scan_gather: null buffer");
    if (n > s->max_elems)
        return fail(c, SFK_E_CAPACITY, "offsets undersized: scanned %u elements, capacity %u", n,
                    s->max_elems);
    SFK_PROF(c, "scan_gather");
    SFK_HIP(c, launch_scan(c->stream, gids, counts, offsets, n, s->sums));
    return SFK_OK;
}

extern "C" int sfk_exclusive_scan(sfk_ctx* c, uint32_t* buf, uint32_t n, sfk_scan_scratch* s) {
    if (!c || !s) return SFK_E_ARG;
    if (n == 0) return SFK_OK;
    if (!buf) return SFK_E_ARG;
    if (n > s->max_elems) return fail(c, SFK_E_CAPACITY, "scan scratch undersized");
    SFK_HIP(c, launch_scan(c->stream, nullptr, buf, buf, n, s->sums));
    return SFK_OK;
}

extern "C" int sfk_argsort_u32_segmented(sfk_ctx* c, sfk_radix_scratch* s, const uint32_t* keys,
                                         const uint32_t* vals, uint32_t nseg, uint32_t seg_len,
                                         uint32_t bits, int write_keys, uint32_t** out_keys,
                                         uint32_t** out_vals) {
    return argsort_impl<uint32_t>(c, s, keys, vals, seg_len, nullptr, bits, write_keys, out_keys,
                                  out_vals, nseg);
}
