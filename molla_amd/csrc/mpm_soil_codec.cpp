// "MOLLSOIL" soil particle checkpoint codec (host only).
//
// Byte format of MpmSoilCheckpoint (crates/molla-solvers/src/mpm_checkpoint.rs):
// header :78-85, digest :87-94, to_bytes :185-214, from_bytes :216-267,
// validate :281-330, rheology_words/validate_rheology :544-592,
// config_words :594-619, Words reader :621-700; the lattice window check is
// validate_lattice_window (mpm_soil_lattice_window.rs:6-39) with
// validate_node_window (soil_candidates_gpu.rs:68-94).  The format is shared
// with the reference so that checkpoints cross between implementations, which
// is why the word order, limits and messages follow it.
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/molla_mpm.h"
#include "blake3.hpp"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the checkpoint payload is copied as little-endian words");

namespace {

constexpr uint32_t VERSION = 1, LATTICE_VERSION = 2;
constexpr uint64_t HEADER = MPM_SOIL_HEADER_BYTES;
constexpr uint32_t WPP = MPM_SOIL_WORDS_PER_PARTICLE;
constexpr uint32_t CFGW = MPM_SOIL_CONFIG_WORDS;
constexpr char MAGIC[8] = {'M', 'O', 'L', 'L', 'S', 'O', 'I', 'L'};
// Words per particle of each stream, and the offset of each stream (in units
// of `count` words): position, velocity, C, F, volume, mass, log-Jp,
// material id, tire compaction.
constexpr uint32_t STRIDES[9] = {4, 4, 12, 12, 1, 1, 1, 1, 4};
constexpr uint32_t STREAM_AT[9] = {0, 4, 8, 20, 32, 33, 34, 35, 36};

thread_local std::string g_err;

int fail(const char* msg) {
    g_err = std::string("MPM soil checkpoint: ") + msg;
    return MPM_ERR_INVALID;
}

inline float f32(uint32_t w) {
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

// config_words indices
enum {
    C_GRID = 0, C_MAXP = 1, C_MAXR = 2, C_H = 3, C_ORIGIN = 4, C_GRAV = 7, C_E = 8, C_NU = 9, C_PHI = 10,
    C_COH = 11, C_WALL = 12, C_MASS = 13, C_VOL = 14, C_RHEO = 15, C_TS = 19, C_ITERS = 20
};

uint64_t max_encoded_bytes(uint32_t max_particles) {
    return HEADER + (23 + 4ull * MPM_MAX_MATERIALS + uint64_t(max_particles) * WPP) * 4;
}

bool window_zero(const int32_t w[3]) { return w[0] == 0 && w[1] == 0 && w[2] == 0; }

void digest(uint32_t version, const uint8_t* payload, uint64_t n, uint8_t out[32]) {
    uint8_t v[4];
    for (int k = 0; k < 4; ++k) v[k] = uint8_t(version >> (8 * k));
    mpm::blake3_hash2(v, version == LATTICE_VERSION ? 4 : 0, payload, n, out);
}

// Words::rheology: canonical encodings only (unused parameter words zero).
bool rheology_canonical(const uint32_t r[4]) {
    switch (r[0]) {
        case 0: return r[2] == 0 && r[3] == 0;
        case 1: return true;
        case 2: return r[3] == 0;
        default: return false;
    }
}

// validate_rheology on a canonical encoding.
bool rheology_valid(const uint32_t r[4]) {
    if (!rheology_canonical(r)) return false;
    const float a = f32(r[1]), b = f32(r[2]), c = f32(r[3]);
    switch (r[0]) {
        case 0: return std::isfinite(a) && a >= 0.0f && a < 1.57079632679489661923f;  // [0, pi/2) in f32
        case 1: return std::isfinite(a) && a >= 0.0f && a < 1.0f && std::isfinite(b) && b >= 0.0f &&
                       std::isfinite(c) && c >= 0.0f;
        default: return std::isfinite(a) && a >= 0.0f && std::isfinite(b) && b >= 0.0f;
    }
}

// Words::config time-step rule: (0, 0) explicit, (1, any) implicit.
bool time_step_canonical(const uint32_t* c) { return (c[C_TS] == 0 && c[C_ITERS] == 0) || c[C_TS] == 1; }

uint64_t payload_words(const mpm_soil_meta* m) {
    return (window_zero(m->grid_window_cells) ? 0 : 3) + CFGW + 1 + 4ull * m->n_materials + 1 +
           uint64_t(m->count) * WPP;
}

struct Reader {
    const uint8_t* p;
    uint64_t left;  // words
    bool next(uint32_t* w) {
        if (left == 0) return false;
        std::memcpy(w, p, 4);
        p += 4;
        --left;
        return true;
    }
};

// from_bytes without validate() and without copying the particle words;
// leaves `words_at` at the first particle word.
int decode_meta(const uint8_t* bytes, uint64_t len, uint32_t max_particles, mpm_soil_meta* m,
                const uint8_t** words_at) {
    if (!bytes && len) return fail("bytes must not be NULL");
    if (!m) return fail("meta must not be NULL");
    if (len < HEADER) return fail("truncated checkpoint header");
    uint32_t version;
    std::memcpy(&version, bytes + 8, 4);
    const bool known = version == VERSION || version == LATTICE_VERSION;
    bool ok = known && len <= max_encoded_bytes(max_particles) + (version == LATTICE_VERSION ? 12 : 0) &&
              std::memcmp(bytes, MAGIC, 8) == 0 && (len - HEADER) % 4 == 0;
    if (ok) {
        uint8_t d[32];
        digest(version, bytes + HEADER, len - HEADER, d);
        ok = std::memcmp(d, bytes + 12, 32) == 0;
    }
    if (!ok) return fail("invalid format, digest or allocation limit");
    std::memset(m, 0, sizeof(*m));
    Reader r{bytes + HEADER, (len - HEADER) / 4};
    if (version == LATTICE_VERSION) {
        for (int a = 0; a < 3; ++a) {
            uint32_t w;
            if (!r.next(&w)) return fail("truncated metadata");
            m->grid_window_cells[a] = int32_t(w);
        }
        if (window_zero(m->grid_window_cells)) return fail("zero-offset checkpoints use version 1");
    }
    for (uint32_t k = 0; k < CFGW; ++k) {
        if (!r.next(&m->config[k])) return fail("truncated metadata");
        if (k == C_RHEO + 3 && !rheology_canonical(&m->config[C_RHEO]))
            return fail("unsupported or noncanonical rheology encoding");
    }
    if (!time_step_canonical(m->config)) return fail("unsupported time step encoding");
    uint32_t rows;
    if (!r.next(&rows)) return fail("truncated metadata");
    if (rows > MPM_MAX_MATERIALS) return fail("material table exceeds capacity");
    m->n_materials = rows;
    for (uint32_t i = 0; i < rows; ++i) {
        for (int k = 0; k < 4; ++k)
            if (!r.next(&m->materials[i][k])) return fail("truncated metadata");
        if (!rheology_canonical(m->materials[i])) return fail("unsupported or noncanonical rheology encoding");
    }
    if (!r.next(&m->count)) return fail("truncated metadata");
    if (m->count > max_particles || r.left != uint64_t(m->count) * WPP)
        return fail("particle count or payload length mismatch");
    if (words_at) *words_at = r.p;
    return MPM_OK;
}

int validate(const mpm_soil_meta* m, const uint32_t* words) {
    const uint32_t* c = m->config;
    const uint64_t g = c[C_GRID];
    const float h = f32(c[C_H]), grav = f32(c[C_GRAV]), E = f32(c[C_E]), nu = f32(c[C_NU]),
                phi = f32(c[C_PHI]), coh = f32(c[C_COH]), wall = f32(c[C_WALL]), pm = f32(c[C_MASS]),
                pv = f32(c[C_VOL]);
    bool origin_ok = true;
    for (int a = 0; a < 3; ++a) origin_ok = origin_ok && std::isfinite(f32(c[C_ORIGIN + a]));
    if (g == 0 || g * g * g > 0xffffffffull || c[C_MAXP] == 0 || c[C_MAXR] == 0 || m->count > c[C_MAXP] ||
        !std::isfinite(h) || h <= 0.0f || !origin_ok || !std::isfinite(grav) || !std::isfinite(E) || E < 0.0f ||
        !std::isfinite(nu) || nu <= -1.0f || nu >= 0.5f || !std::isfinite(phi) || !std::isfinite(coh) ||
        coh < 0.0f || !std::isfinite(wall) || !(wall >= 0.0f && wall <= 1.0f) || !std::isfinite(pm) || pm <= 0.0f ||
        !std::isfinite(pv) || pv <= 0.0f || m->n_materials > MPM_MAX_MATERIALS || !time_step_canonical(c))
        return fail("invalid configuration or state length");
    if (m->count && !words) return fail("invalid configuration or state length");
    if (!rheology_valid(&c[C_RHEO])) return fail("invalid material parameters");
    for (uint32_t i = 0; i < m->n_materials; ++i)
        if (!rheology_valid(m->materials[i])) return fail("invalid material parameters");
    const uint64_t n = m->count;
    const uint32_t rows = m->n_materials ? m->n_materials : 1;
    for (int s = 0; s < 9; ++s) {
        const uint32_t* w = words + STREAM_AT[s] * n;
        const uint64_t len = uint64_t(STRIDES[s]) * n;
        if (s == 7) {
            for (uint64_t i = 0; i < len; ++i)
                if (w[i] >= rows) return fail("material ID outside authoritative table");
            continue;
        }
        for (uint64_t i = 0; i < len; ++i) {
            const float v = f32(w[i]);
            if (!std::isfinite(v) || ((s == 4 || s == 5) && v <= 0.0f))
                return fail("non-finite history or nonpositive mass/reference volume");
        }
    }
    return MPM_OK;
}

}  // namespace

extern "C" {

const char* mpm_soil_last_error(void) { return g_err.c_str(); }

int mpm_blake3(const void* data, uint64_t len, uint8_t out[32]) {
    if ((!data && len) || !out) return fail("NULL argument");
    mpm::blake3_hash(data, len, out);
    return MPM_OK;
}

uint64_t mpm_soil_encoded_bytes(const mpm_soil_meta* meta) {
    return meta ? HEADER + 4 * payload_words(meta) : 0;
}

int mpm_soil_encode(const mpm_soil_meta* m, const uint32_t* words, uint8_t* out, uint64_t cap) {
    if (!m || !out) return fail("NULL argument");
    if (m->n_materials > MPM_MAX_MATERIALS) return fail("material table exceeds capacity");
    if (m->count && !words) return fail("words must not be NULL");
    const uint64_t total = mpm_soil_encoded_bytes(m);
    if (cap < total) return fail("output buffer too small");
    const uint32_t version = window_zero(m->grid_window_cells) ? VERSION : LATTICE_VERSION;
    std::memcpy(out, MAGIC, 8);
    std::memcpy(out + 8, &version, 4);
    uint8_t* p = out + HEADER;
    auto put = [&p](uint32_t w) {
        std::memcpy(p, &w, 4);
        p += 4;
    };
    if (version == LATTICE_VERSION)
        for (int a = 0; a < 3; ++a) put(uint32_t(m->grid_window_cells[a]));
    for (uint32_t k = 0; k < CFGW; ++k) put(m->config[k]);
    put(m->n_materials);
    for (uint32_t i = 0; i < m->n_materials; ++i)
        for (int k = 0; k < 4; ++k) put(m->materials[i][k]);
    put(m->count);
    if (m->count) std::memcpy(p, words, uint64_t(m->count) * WPP * 4);
    uint8_t d[32];
    digest(version, out + HEADER, total - HEADER, d);
    std::memcpy(out + 12, d, 32);
    return MPM_OK;
}

int mpm_soil_decode_meta(const uint8_t* bytes, uint64_t len, uint32_t max_particles, mpm_soil_meta* meta) {
    return decode_meta(bytes, len, max_particles, meta, nullptr);
}

int mpm_soil_decode(const uint8_t* bytes, uint64_t len, uint32_t max_particles, mpm_soil_meta* meta,
                    uint32_t* words, uint64_t words_cap) {
    const uint8_t* at = nullptr;
    int rc = decode_meta(bytes, len, max_particles, meta, &at);
    if (rc != MPM_OK) return rc;
    const uint64_t need = uint64_t(meta->count) * WPP;
    if (need && (!words || words_cap < need)) return fail("word buffer too small");
    if (need) std::memcpy(words, at, need * 4);
    return validate(meta, words);
}

int mpm_soil_validate(const mpm_soil_meta* meta, const uint32_t* words) {
    if (!meta) return fail("meta must not be NULL");
    return validate(meta, words);
}

int mpm_soil_validate_window(const mpm_soil_meta* m, const uint32_t* words) {
    if (!m) return fail("meta must not be NULL");
    if (m->count && !words) return fail("words must not be NULL");
    const int64_t limit = (int64_t(1) << 23) - 1;
    const uint32_t size = m->config[C_GRID];
    const float h = f32(m->config[C_H]);
    if (size == 0 || !std::isfinite(h) || h <= 0.0f) return fail("invalid node window spacing or extent");
    for (int a = 0; a < 3; ++a) {
        const float o = f32(m->config[C_ORIGIN + a]);
        const int64_t lo = m->grid_window_cells[a], hi = lo + int64_t(size) - 1;
        if (lo < -limit || hi > limit || !std::isfinite(o) || !std::isfinite(o + float(lo) * h) ||
            !std::isfinite(o + float(hi) * h))
            return fail("node window exceeds finite half-cell coordinate range");
    }
    const float inv_h = 1.0f / h;
    if (!std::isfinite(inv_h)) return fail("integer window requires finite inverse cell size");
    for (uint64_t p = 0; p < m->count; ++p)
        for (int a = 0; a < 3; ++a) {
            const float xp = (f32(words[4 * p + a]) - f32(m->config[C_ORIGIN + a])) * inv_h;
            const float base = std::floor(xp - 0.5f);
            if (!std::isfinite(base) || double(base) < -double(limit) || double(base) + 2.0 > double(limit))
                return fail("particle stencil exceeds exact lattice index range");
            const int64_t local = int64_t(base) - int64_t(m->grid_window_cells[a]);
            if (local < 0 || local + 2 >= int64_t(size)) return fail("particle stencil is outside the integer grid window");
        }
    return MPM_OK;
}

}  // extern "C"
