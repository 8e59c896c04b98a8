// Persisted transformed weights: U written to a file once and loaded in later processes,
// keyed by the weights' hash and the packed-layout version.
//
// The reference computes U in a separate transform_filter buffer (conv2d.h:14, main.cpp:38)
// on every run; a loader that trusts a stale file would return wrong convolutions, so a
// load re-hashes the caller's current weights and refuses any mismatch.
//
// File layout (little-endian, 64-byte header, then the payload):
//   0  char[8]  magic "WINOU16\0"
//   8  u32      layout version (CONV2D_U_LAYOUT_VERSION)
//  12  u32      U layout: 0 Winograd F(2x2) (CONV2D_METHOD_WINOGRAD), 1 direct path
//               (CONV2D_METHOD_DIRECT), 2 Winograd F(4x4) (the problem's form is CONV2D_FORM_F43)
//  16  u32 x4   k, c, r, s
//  32  u64      FNV-1a 64 of the K*C*R*S FP16 weights
//  40  u64      payload bytes (the problem's U size)
//  48  u64      FNV-1a 64 of the payload
//  56  u64      zero
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "conv2d.h"
#include "plugin_internal.h"

namespace {

constexpr char kMagic[8] = {'W', 'I', 'N', 'O', 'U', '1', '6', '\0'};
constexpr uint32_t kLayoutF43 = 2;  // header layout value of the F(4x4) U

struct Header {
    char magic[8];
    uint32_t version, method;
    uint32_t k, c, r, s;
    uint64_t weight_hash, payload_bytes, payload_hash, zero;
};
static_assert(sizeof(Header) == 64, "file header is 64 bytes");

uint64_t fnv1a64(const void* data, size_t n) {
    const unsigned char* p = (const unsigned char*)data;
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

// Device bytes -> host vector on the caller's stream (synchronised).
int copy_to_host(const void* dev, size_t n, std::vector<unsigned char>* out, hipStream_t st) {
    out->resize(n);
    if (hipMemcpyAsync(out->data(), dev, n, hipMemcpyDeviceToHost, st) != hipSuccess) return CONV2D_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return CONV2D_ERR_HIP;
    return CONV2D_OK;
}

int problem_header(const problem_t* pb, Header* h, size_t* ubytes) {
    wino::Geometry g;
    int rc = wino::derive_geometry(pb, &g);
    if (rc) return rc;
    memset(h, 0, sizeof(*h));
    memcpy(h->magic, kMagic, sizeof(kMagic));
    h->version = CONV2D_U_LAYOUT_VERSION;
    const bool f43 = wino::select_form(pb, g) == CONV2D_FORM_F43;
    h->method = g.direct ? CONV2D_METHOD_DIRECT : (f43 ? kLayoutF43 : CONV2D_METHOD_WINOGRAD);
    h->k = pb->k;
    h->c = pb->c;
    h->r = pb->r;
    h->s = pb->s;
    h->payload_bytes = f43 ? g.u43_bytes : g.u_bytes;
    *ubytes = h->payload_bytes;
    return CONV2D_OK;
}

size_t weight_bytes(const problem_t* pb) { return (size_t)pb->k * pb->c * pb->r * pb->s * 2; }

}  // namespace

int conv2d_weight_hash(const void* host_bytes, size_t nbytes, unsigned long long* hash) {
    if (!hash || (!host_bytes && nbytes)) return CONV2D_ERR_NULL;
    *hash = fnv1a64(host_bytes, nbytes);
    return CONV2D_OK;
}

int conv2d_weights_save(const problem_t* pb, const char* path, void* stream) {
    if (!pb || !path || !pb->weight || !pb->transform_filter) return CONV2D_ERR_NULL;
    Header h;
    size_t ubytes;
    int rc = problem_header(pb, &h, &ubytes);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    std::vector<unsigned char> w, u;
    if ((rc = copy_to_host(pb->weight, weight_bytes(pb), &w, st))) return rc;
    if ((rc = copy_to_host(pb->transform_filter, ubytes, &u, st))) return rc;
    h.weight_hash = fnv1a64(w.data(), w.size());
    h.payload_hash = fnv1a64(u.data(), u.size());
    FILE* f = fopen(path, "wb");
    if (!f) return CONV2D_ERR_IO;
    const bool ok = fwrite(&h, sizeof(h), 1, f) == 1 && fwrite(u.data(), 1, u.size(), f) == u.size();
    if (fclose(f) != 0 || !ok) return CONV2D_ERR_IO;
    return CONV2D_OK;
}

int conv2d_weights_load(const problem_t* pb, const char* path, void* stream) {
    if (!pb || !path || !pb->weight || !pb->transform_filter) return CONV2D_ERR_NULL;
    Header want;
    size_t ubytes;
    int rc = problem_header(pb, &want, &ubytes);
    if (rc) return rc;
    FILE* f = fopen(path, "rb");
    if (!f) return CONV2D_ERR_IO;
    Header got;
    if (fread(&got, sizeof(got), 1, f) != 1) {
        fclose(f);
        return CONV2D_ERR_IO;
    }
    // everything the header can refute, before any device access
    if (memcmp(got.magic, kMagic, sizeof(kMagic)) != 0 || got.version != want.version || got.method != want.method ||
        got.k != want.k || got.c != want.c || got.r != want.r || got.s != want.s ||
        got.payload_bytes != want.payload_bytes || got.zero != 0) {
        fclose(f);
        return CONV2D_ERR_MISMATCH;
    }
    std::vector<unsigned char> u(ubytes);
    const bool full = fread(u.data(), 1, ubytes, f) == ubytes;
    const bool trailing = fgetc(f) != EOF;
    fclose(f);
    if (!full) return CONV2D_ERR_IO;
    if (trailing || fnv1a64(u.data(), u.size()) != got.payload_hash) return CONV2D_ERR_MISMATCH;
    hipStream_t st = (hipStream_t)stream;
    std::vector<unsigned char> w;
    if ((rc = copy_to_host(pb->weight, weight_bytes(pb), &w, st))) return rc;
    if (fnv1a64(w.data(), w.size()) != got.weight_hash) return CONV2D_ERR_MISMATCH;
    if (hipMemcpyAsync(pb->transform_filter, u.data(), ubytes, hipMemcpyHostToDevice, st) != hipSuccess)
        return CONV2D_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return CONV2D_ERR_HIP;
    return CONV2D_OK;
}
