// state.hip -- checkpoint and restart of a context: everything the next step
// reads, copied to or from a caller-owned host blob, so that a run restarted
// from a saved state continues bit for bit.  This is the role of the
// reference's backupData (seq/bouyancy3d.F90:1011-1030: u, v, w, T, f, g to
// 'backupFile-<itc>.bin') and its loadInitField switch
// (3d/mpi_blocked/initial.F90:102, 195-206); the blob holds the stored
// post-collision populations in the device layout rather than the
// reference's pre-collision f, because that is what the fused step resumes
// from.
//
// Blob layout: a fixed header (StateHeader), then in order
//   f        Q * ps elements of the storage type (current buffer)
//   lid_rho  nx * ny doubles
//   macro    4 * n doubles (rho, u, v, w)
//   prev     2 * n doubles (up, vp of the residual)
//   halo     htotal elements (incoming halo regions), when halo faces are set
//   thermal  g (Q7 * ps elements), force (3 * ps, FP64 or FP32), T (n), wp/Tp (2 * n)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "context.hpp"
#include "lattice.hpp"

namespace {

constexpr char STATE_MAGIC[8] = {'M', 'G', 'L', 'C', 'S', 'T', 'A', '2'};

// What a blob's populations mean: besides the grid and storage, FP16 blobs
// hold (f - w) * 2^scale_log2, and the halo flags decide which halo the
// next step reads and whether it writes one.  A context must match all of
// them to resume.
struct StateHeader {
    char magic[8];
    int32_t header_bytes;
    int32_t nx, ny, nz;
    int32_t storage, collision, thermal, halo_faces;
    int32_t scale_log2;     // the FP16 scale exponent in effect (0 for FP64 / FP32)
    int32_t halo_loopback, halo_overlap;
    int32_t reserved;
    int64_t ps, htotal, itc;
    int64_t total_bytes;
};

// The FP16 scale exponent in effect for c (f16_scale_log2 = 0 selects 12).
int32_t scale_log2_of(const lbm_ctx* c) {
    if (c->p.storage != LBM_STORE_F16) return 0;
    return c->p.f16_scale_log2 ? c->p.f16_scale_log2 : 12;
}

struct Section {
    void* dev;
    size_t bytes;
};

// The device arrays of the state, in blob order.
int sections(lbm_ctx* c, Section* out) {
    int k = 0;
    const size_t n = (size_t)c->n;
    out[k++] = {c->buf[c->cur], c->elem * (size_t)c->ps * lbm::Q};
    out[k++] = {c->lid_rho, sizeof(double) * (size_t)c->p.nx * c->p.ny};
    out[k++] = {c->macro, sizeof(double) * 4 * n};
    out[k++] = {c->prev, sizeof(double) * 2 * n};
    if (c->p.halo_faces) out[k++] = {c->hin, c->elem * (size_t)c->htotal};
    if (c->p.thermal) {
        out[k++] = {c->gbuf[c->cur], c->elem * (size_t)c->ps * lbm::Q7};
        out[k++] = {c->force, thermal_force_elem(c) * 3 * (size_t)c->ps};
        out[k++] = {c->temp, sizeof(double) * n};
        out[k++] = {c->prev_t, sizeof(double) * 2 * n};
    }
    return k;
}

size_t blob_bytes(lbm_ctx* c) {
    Section s[9];
    const int k = sections(c, s);
    size_t total = sizeof(StateHeader);
    for (int q = 0; q < k; ++q) total += s[q].bytes;
    return total;
}

int usable(lbm_ctx* c, const char* who) {
    if (!c) return lbm_fail(LBM_ERR_ARG, "%s: null context", who);
    if (c->part)
        return lbm_fail(LBM_ERR_UNSUPPORTED, "%s: contexts with attached spheres are not checkpointed", who);
    if (c->in_step) return lbm_fail(LBM_ERR_STATE, "%s between lbm_step_begin and lbm_step_end", who);
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_state_bytes(lbm_ctx* c, size_t* bytes) {
    int rc = usable(c, "lbm_state_bytes");
    if (rc) return rc;
    if (!bytes) return lbm_fail(LBM_ERR_ARG, "lbm_state_bytes: null argument");
    *bytes = blob_bytes(c);
    return LBM_OK;
}

int lbm_save_state(lbm_ctx* c, void* blob, size_t bytes) {
    int rc = usable(c, "lbm_save_state");
    if (rc) return rc;
    if (!blob) return lbm_fail(LBM_ERR_ARG, "lbm_save_state: null buffer");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_save_state before lbm_initial / lbm_load_fields");
    const size_t need = blob_bytes(c);
    if (bytes < need) return lbm_fail(LBM_ERR_ARG, "lbm_save_state: buffer of %zu bytes, state needs %zu", bytes, need);
    StateHeader h{};
    std::memcpy(h.magic, STATE_MAGIC, sizeof h.magic);
    h.header_bytes = (int32_t)sizeof(StateHeader);
    h.nx = c->p.nx;
    h.ny = c->p.ny;
    h.nz = c->p.nz;
    h.storage = c->p.storage;
    h.collision = c->p.collision;
    h.thermal = c->p.thermal;
    h.halo_faces = c->p.halo_faces;
    h.scale_log2 = scale_log2_of(c);
    h.halo_loopback = c->p.halo_loopback;
    h.halo_overlap = c->p.halo_overlap;
    h.ps = c->ps;
    h.htotal = c->htotal;
    h.itc = c->itc;
    h.total_bytes = (int64_t)need;
    char* out = static_cast<char*>(blob);
    std::memcpy(out, &h, sizeof h);
    size_t off = sizeof h;
    Section s[9];
    const int k = sections(c, s);
    for (int q = 0; q < k; ++q) {
        HIP_TRY(hipMemcpyAsync(out + off, s[q].dev, s[q].bytes, hipMemcpyDeviceToHost, c->stream));
        off += s[q].bytes;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_load_state(lbm_ctx* c, const void* blob, size_t bytes) {
    int rc = usable(c, "lbm_load_state");
    if (rc) return rc;
    if (!blob || bytes < sizeof(StateHeader)) return lbm_fail(LBM_ERR_ARG, "lbm_load_state: no state in buffer");
    StateHeader h;
    std::memcpy(&h, blob, sizeof h);
    if (std::memcmp(h.magic, STATE_MAGIC, sizeof h.magic) != 0 || h.header_bytes != (int32_t)sizeof(StateHeader))
        return lbm_fail(LBM_ERR_ARG, "lbm_load_state: not a state written by lbm_save_state");
    if (h.nx != c->p.nx || h.ny != c->p.ny || h.nz != c->p.nz || h.storage != c->p.storage ||
        h.collision != c->p.collision || h.thermal != c->p.thermal || h.halo_faces != c->p.halo_faces ||
        h.ps != c->ps || h.htotal != c->htotal)
        return lbm_fail(LBM_ERR_ARG,
                        "lbm_load_state: state of a %dx%dx%d storage %d collision %d thermal %d halo 0x%x "
                        "context, this one is %dx%dx%d storage %d collision %d thermal %d halo 0x%x",
                        h.nx, h.ny, h.nz, h.storage, h.collision, h.thermal, h.halo_faces, c->p.nx, c->p.ny,
                        c->p.nz, c->p.storage, c->p.collision, c->p.thermal, c->p.halo_faces);
    if (h.scale_log2 != scale_log2_of(c))
        return lbm_fail(LBM_ERR_ARG,
                        "lbm_load_state: FP16 state scaled by 2^%d, this context scales by 2^%d "
                        "(f16_scale_log2)", h.scale_log2, scale_log2_of(c));
    if (h.halo_loopback != c->p.halo_loopback || h.halo_overlap != c->p.halo_overlap)
        return lbm_fail(LBM_ERR_ARG,
                        "lbm_load_state: state of a context with halo_loopback %d halo_overlap %d, this one "
                        "has %d %d", h.halo_loopback, h.halo_overlap, c->p.halo_loopback, c->p.halo_overlap);
    const size_t need = blob_bytes(c);
    if (h.total_bytes != (int64_t)need || bytes < need)
        return lbm_fail(LBM_ERR_ARG, "lbm_load_state: state of %lld bytes, %zu given, %zu expected",
                        (long long)h.total_bytes, bytes, need);
    c->cur = 0;  // the saved current buffer goes to buffer 0
    const char* in = static_cast<const char*>(blob);
    size_t off = sizeof h;
    Section s[9];
    const int k = sections(c, s);
    for (int q = 0; q < k; ++q) {
        HIP_TRY(hipMemcpyAsync(s[q].dev, in + off, s[q].bytes, hipMemcpyHostToDevice, c->stream));
        off += s[q].bytes;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));  // the caller may free the blob on return
    c->itc = h.itc;
    c->initialised = true;
    c->halo_fresh = false;  // the next looped-back split step exchanges the loaded state's halo first
    return LBM_OK;
}

}  // extern "C"
