// State checkpoint / restart of one context (include/capaz.h:
// capaz_save_state, capaz_load_state).
//
// The reference's save/restore (main.cpp:1035-1145) writes S_w, P_w, S_n,
// roS_*_old and m after a 4-byte step counter; it leaves out the boundary
// values of P_n and P_g that persist between steps, saves roS_g_old from the
// roS_n_old device array, and its restore reads Nx*Ny*Ny doubles per field.
// This format keeps every field the next step reads, so a restored context
// continues bit for bit:
//
//   8 bytes   "CAPAZST1"
//   int32 x 6 version (1), model, energy, Nx, Ny, Nz
//   int64     step
//   int32     nfields, then nfields field ids (capaz_field_id)
//   nfields x Nx*Ny*Nz doubles, x fastest, in the listed order
//
// all little-endian as the host writes it.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capaz_ctx.h"

namespace {

using capaz::detail::fail;

const char kMagic[8] = {'C', 'A', 'P', 'A', 'Z', 'S', 'T', '1'};

// The fields a step reads: the primary state and the porosity / permeability.
std::vector<int32_t> state_fields(const capaz_ctx* c) {
    std::vector<int32_t> f = {capaz::F_P_w, capaz::F_S_n, capaz::F_m, capaz::F_K};
    if (c->user.model == 3) {
        f.push_back(capaz::F_S_w);
        f.push_back(capaz::F_P_n);
        f.push_back(capaz::F_P_g);
    }
    if (c->user.energy) f.push_back(capaz::F_T);
    return f;
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int capaz_save_state(capaz_ctx* c, const char* path, long step) {
    if (!c || !path) return CAPAZ_ERR_INVALID;
    const std::vector<int32_t> ids = state_fields(c);
    const size_t n = (size_t)c->p.n;
    std::vector<double> buf(n);
    File out;
    if (!(out.f = std::fopen(path, "wb"))) return fail(c, CAPAZ_ERR_INVALID, std::string("cannot open ") + path);
    const int32_t head[6] = {1, c->user.model, c->user.energy, c->user.Nx, c->user.Ny, c->user.Nz};
    const int64_t st = step;
    const int32_t nf = (int32_t)ids.size();
    bool ok = std::fwrite(kMagic, 1, 8, out.f) == 8 && std::fwrite(head, sizeof(head), 1, out.f) == 1 &&
              std::fwrite(&st, sizeof(st), 1, out.f) == 1 && std::fwrite(&nf, sizeof(nf), 1, out.f) == 1 &&
              std::fwrite(ids.data(), sizeof(int32_t), ids.size(), out.f) == ids.size();
    for (size_t q = 0; ok && q < ids.size(); ++q) {
        int rc = capaz_get_field(c, ids[q], buf.data(), n);
        if (rc) return rc;
        ok = std::fwrite(buf.data(), sizeof(double), n, out.f) == n;
    }
    if (!ok || std::fflush(out.f) != 0) return fail(c, CAPAZ_ERR_INVALID, std::string("write failed: ") + path);
    return CAPAZ_OK;
}

int capaz_load_state(capaz_ctx* c, const char* path, long* step) {
    if (!c || !path) return CAPAZ_ERR_INVALID;
    File in;
    if (!(in.f = std::fopen(path, "rb"))) return fail(c, CAPAZ_ERR_INVALID, std::string("cannot open ") + path);
    char magic[8];
    int32_t head[6];
    int64_t st = 0;
    int32_t nf = 0;
    if (std::fread(magic, 1, 8, in.f) != 8 || std::memcmp(magic, kMagic, 8) != 0 ||
        std::fread(head, sizeof(head), 1, in.f) != 1 || std::fread(&st, sizeof(st), 1, in.f) != 1 ||
        std::fread(&nf, sizeof(nf), 1, in.f) != 1)
        return fail(c, CAPAZ_ERR_INVALID, std::string("not a capaz state file: ") + path);
    if (head[0] != 1 || head[1] != c->user.model || head[2] != c->user.energy || head[3] != c->user.Nx ||
        head[4] != c->user.Ny || head[5] != c->user.Nz)
        return fail(c, CAPAZ_ERR_INVALID, "state file: version, model or grid differs from the context");
    if (nf < 1 || nf > capaz::F_COUNT) return fail(c, CAPAZ_ERR_INVALID, "state file: bad field count");
    std::vector<int32_t> ids((size_t)nf);
    if (std::fread(ids.data(), sizeof(int32_t), ids.size(), in.f) != ids.size())
        return fail(c, CAPAZ_ERR_INVALID, "state file: truncated");
    const size_t n = (size_t)c->p.n;
    std::vector<double> buf(n);
    for (int32_t id : ids) {
        if (!capaz_has_field(c, id)) return fail(c, CAPAZ_ERR_INVALID, "state file: field not in this model");
        if (std::fread(buf.data(), sizeof(double), n, in.f) != n)
            return fail(c, CAPAZ_ERR_INVALID, "state file: truncated");
        int rc = capaz_set_field(c, id, buf.data(), n);
        if (rc) return rc;
    }
    if (step) *step = (long)st;
    return CAPAZ_OK;
}

}  // extern "C"
