// Scene ingestion from binary little-endian 3DGS PLY (src/ply.rs:9-112).
// The host reads only the header; the vertex rows go to the device as they
// lie in the file and one kernel transposes them into the field-major
// attribute and SH planes (src/layout.rs:41-61), normalising the rotation
// on the way.  A workgroup stages 64 rows in LDS with coalesced loads, then
// writes every plane's 64 values as one contiguous run.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "sfk_device.h"
#include "sfk_internal.h"

namespace sfk {

namespace {

constexpr uint32_t kPlyRows = 64;

struct PlyArgs {
    const float* rows;
    float* attrs;
    float* sh;
    uint64_t n;
    uint32_t vec4;  // rows 16-byte aligned: stage with float4 loads
    uint32_t stride, n_rest, sh_planes;
    int32_t base[14];
    int32_t rest[SFK_PLY_MAX_REST];
};

// Columns of PlyArgs::base, in the reference's lookup order (ply.rs:47-48).
enum { kX, kY, kZ, kS0, kS1, kS2, kOp, kR0, kR1, kR2, kR3, kDc0, kDc1, kDc2 };

__global__ __launch_bounds__(256) void k_ply_decode(PlyArgs a) {
    __shared__ __align__(16) float rows[kPlyRows * SFK_PLY_MAX_STRIDE];
    const uint64_t v0 = (uint64_t)blockIdx.x * kPlyRows;
    const uint32_t nv = (uint32_t)min((uint64_t)kPlyRows, a.n - v0);
    // The block's rows start 64 * stride floats apart: with a 16-byte
    // aligned base they load as float4 (a ragged last block finishes with
    // scalars); rows at an arbitrary float offset of the file load singly.
    const float* src = a.rows + v0 * a.stride;
    const uint32_t total = nv * a.stride, n4 = a.vec4 ? total / 4 : 0u;
    const float4* src4 = reinterpret_cast<const float4*>(src);
    float4* rows4 = reinterpret_cast<float4*>(rows);
    for (uint32_t e = threadIdx.x; e < n4; e += 256) rows4[e] = src4[e];
    for (uint32_t e = 4 * n4 + threadIdx.x; e < total; e += 256) rows[e] = src[e];
    __syncthreads();
    const uint32_t per_ch = a.n_rest / 3;
    const uint32_t planes = SFK_ATTR_PLANES + a.sh_planes;
    for (uint32_t e = threadIdx.x; e < planes * kPlyRows; e += 256) {
        const uint32_t p = e / kPlyRows, v = e % kPlyRows;
        if (v >= nv) continue;
        const float* r = rows + v * a.stride;
        float out;
        if (p < 3) {
            out = r[a.base[kX + p]];
        } else if (p < 7) {
            // glam Quat::from_xyzw(rot_1, rot_2, rot_3, rot_0).normalize()
            // (ply.rs:84-85), its x86-64 evaluation: the squared length
            // summed as (x^2 + z^2) + (y^2 + w^2), one sqrt, and each
            // component divided by the length.
            const float x = r[a.base[kR1]], y = r[a.base[kR2]], z = r[a.base[kR3]],
                        w = r[a.base[kR0]];
            const float len = sqrtf((x * x + z * z) + (y * y + w * w));
            const float c = p == 3 ? w : p == 4 ? x : p == 5 ? y : z;
            out = c / len;
        } else if (p < 10) {
            out = r[a.base[kS0 + (p - 7)]];
        } else if (p == 10) {
            out = r[a.base[kOp]];
        } else {
            const uint32_t q = p - SFK_ATTR_PLANES;  // SH plane: coefficient q/3, channel q%3
            if (q < 3) {
                out = r[a.base[kDc0 + q]];
            } else {
                const uint32_t j = q / 3 - 1, c = q % 3;
                // f_rest_k is channel-major in the file (k = c * per_ch + j).
                out = j < per_ch ? r[a.rest[c * per_ch + j]] : 0.0f;
            }
        }
        if (p < SFK_ATTR_PLANES)
            a.attrs[(uint64_t)p * a.n + v0 + v] = out;
        else
            a.sh[(uint64_t)(p - SFK_ATTR_PLANES) * a.n + v0 + v] = out;
    }
}

int ply_error(char* err, int err_len, const char* fmt, const std::string& a,
              const std::string& b = std::string()) {
    if (err && err_len > 0) snprintf(err, (size_t)err_len, fmt, a.c_str(), b.c_str());
    return SFK_E_ARG;
}

// Rust's usize parse: optional '+', then decimal digits only.
bool parse_count(const std::string& s, uint64_t* out) {
    size_t i = (!s.empty() && s[0] == '+') ? 1 : 0;
    if (i >= s.size()) return false;
    uint64_t v = 0;
    for (; i < s.size(); ++i) {
        if (s[i] < '0' || s[i] > '9') return false;
        const uint64_t d = (uint64_t)(s[i] - '0');
        if (v > (UINT64_MAX - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

}  // namespace

hipError_t launch_ply_decode(hipStream_t s, const sfk_ply_layout& l, const float* rows,
                             float* attrs, float* sh) {
    if (l.n == 0) return hipSuccess;
    PlyArgs a;
    memset(&a, 0, sizeof a);
    a.rows = rows;
    a.attrs = attrs;
    a.sh = sh;
    a.n = l.n;
    a.stride = l.stride;
    a.n_rest = l.n_rest;
    a.sh_planes = l.sh_planes;
    a.vec4 = (reinterpret_cast<uintptr_t>(rows) & 15u) == 0;
    memcpy(a.base, l.base, sizeof a.base);
    memcpy(a.rest, l.rest, sizeof a.rest);
    const uint64_t blocks = (l.n + kPlyRows - 1) / kPlyRows;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    k_ply_decode<<<(uint32_t)blocks, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace sfk

using namespace sfk;

extern "C" int sfk_ply_parse_header(const void* bytes, size_t len, sfk_ply_layout* out, char* err,
                                    int err_len) {
    if (!bytes || !out) return SFK_E_ARG;
    memset(out, 0, sizeof *out);
    if (err && err_len > 0) err[0] = 0;
    const char* p = static_cast<const char*>(bytes);
    size_t pos = 0, data = len;
    uint64_t count = 0;
    std::vector<std::string> props;
    // Header lines up to end_header (ply.rs:13-32); the rows start right
    // after its newline.
    while (pos < len) {
        size_t e = pos;
        while (e < len && p[e] != '\n') ++e;
        std::vector<std::string> tok;
        size_t i = pos;
        while (i < e) {
            while (i < e && isspace((unsigned char)p[i])) ++i;
            size_t j = i;
            while (j < e && !isspace((unsigned char)p[j])) ++j;
            if (j > i) tok.emplace_back(p + i, j - i);
            i = j;
        }
        pos = e < len ? e + 1 : len;
        if (tok.empty()) continue;
        if (tok[0] == "end_header") {
            data = pos;
            break;
        }
        if (tok[0] == "format" && tok.size() >= 2 && tok[1] != "binary_little_endian")
            return ply_error(err, err_len, "Unsupported PLY format: %s", tok[1]);
        if (tok.size() == 3 && tok[0] == "element" && tok[1] == "vertex") {
            if (!parse_count(tok[2], &count)) return ply_error(err, err_len, "Invalid vertex count%s", "");
        } else if (tok.size() == 3 && tok[0] == "property") {
            if (tok[1] != "float")
                return ply_error(err, err_len, "Unsupported property type %s: %s", tok[1], tok[2]);
            props.push_back(tok[2]);
        }
    }
    if (count == 0) return ply_error(err, err_len, "PLY contains no vertices%s", "");
    static const char* kBase[14] = {"x",       "y",       "z",       "scale_0", "scale_1",
                                    "scale_2", "opacity", "rot_0",   "rot_1",   "rot_2",
                                    "rot_3",   "f_dc_0",  "f_dc_1",  "f_dc_2"};
    for (int k = 0; k < 14; ++k) {
        auto it = std::find(props.begin(), props.end(), kBase[k]);
        if (it == props.end()) return ply_error(err, err_len, "Missing property: %s", kBase[k]);
        out->base[k] = (int32_t)(it - props.begin());
    }
    std::vector<std::pair<uint64_t, int32_t>> rest;  // (number, column)
    for (size_t c = 0; c < props.size(); ++c) {
        const std::string& s = props[c];
        uint64_t k;
        if (s.compare(0, 7, "f_rest_") == 0 && parse_count(s.substr(7), &k))
            rest.emplace_back(k, (int32_t)c);
    }
    std::stable_sort(rest.begin(), rest.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    if (rest.size() > SFK_PLY_MAX_REST)
        return ply_error(err, err_len, "%s f_rest properties (at most 45: SH degree 3)",
                         std::to_string(rest.size()));
    if (props.size() > SFK_PLY_MAX_STRIDE)
        return ply_error(err, err_len, "%s properties per vertex (at most 128)",
                         std::to_string(props.size()));
    out->n = count;
    out->stride = (uint32_t)props.size();
    out->n_rest = (uint32_t)rest.size();
    out->sh_planes = out->n_rest + 3;  // ply.rs:75
    out->sh_per_ch = out->n_rest / 3 + 1;
    out->data_offset = data;
    for (size_t k = 0; k < rest.size(); ++k) out->rest[k] = rest[k].second;
    // Every vertex must be in the file (ply.rs:80-82 read_exact).
    const uint64_t row_bytes = (uint64_t)out->stride * 4;
    if (data > len || (len - data) / row_bytes < count) {
        const uint64_t have = data > len ? 0 : (len - data) / row_bytes;
        return ply_error(err, err_len, "failed to read vertex %s of %s", std::to_string(have),
                         std::to_string(count));
    }
    return SFK_OK;
}

extern "C" int sfk_ply_decode(sfk_ctx* c, const sfk_ply_layout* l, const float* rows,
                              float* attrs, float* sh) {
    if (!c || !l || !rows || !attrs || !sh || l->stride == 0 || l->stride > SFK_PLY_MAX_STRIDE ||
        l->n_rest > SFK_PLY_MAX_REST || l->sh_planes != l->n_rest + 3)
        return SFK_E_ARG;
    for (int k = 0; k < 14; ++k)
        if (l->base[k] < 0 || (uint32_t)l->base[k] >= l->stride)
            return fail(c, SFK_E_ARG, "ply_decode: column %d out of the row", l->base[k]);
    for (uint32_t k = 0; k < l->n_rest; ++k)
        if (l->rest[k] < 0 || (uint32_t)l->rest[k] >= l->stride)
            return fail(c, SFK_E_ARG, "ply_decode: column %d out of the row", l->rest[k]);
    // A device copy of the file plus data_offset lands wherever the header
    // ended: float loads need 4-byte alignment (16 for the vector path,
    // chosen per call).
    if (reinterpret_cast<uintptr_t>(rows) & 3u)
        return fail(c, SFK_E_ARG,
                    "ply_decode: rows at a misaligned address (copy them to a 4-byte boundary)");
    SFK_PROF(c, "ply_decode");
    SFK_HIP(c, launch_ply_decode(c->stream, *l, rows, attrs, sh));
    return SFK_OK;
}
