// capi_state.hip - binary state export / import (ifl_save_state, ifl_load_state).
#include "capi_host.hpp"

using namespace ifl_host;

// ---------------- binary state files ----------------
namespace {

constexpr int kStateFields[8] = {IFL_FIELD_U, IFL_FIELD_V, IFL_FIELD_W, IFL_FIELD_P,
                                 IFL_FIELD_TEMP, IFL_FIELD_UF, IFL_FIELD_VF, IFL_FIELD_WF};
static_assert(sizeof(ifl_state_header) == 160, "ifl_state_header layout");

uint64_t fnv1a(const unsigned char *p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

size_t statePayloadFloats(const ifl_config &c) {
    size_t n = 0;
    for (int f : kStateFields) n += fieldSize(c, f);
    return n;
}

}  // namespace

extern "C" {

int ifl_save_state(ifl_solver *h, const char *path) {
    if (!h || !path) return fail(IFL_ERR_ARG, "null argument");
    const ifl_config &c = h->cfg;
    std::vector<float> data(statePayloadFloats(c));
    size_t off = 0;
    for (int f : kStateFields) {
        const size_t n = fieldSize(c, f);
        IFL_HIP(hipMemcpyAsync(data.data() + off, fieldPtr(h, f), n * 4, hipMemcpyDeviceToHost, h->stream));
        off += n;
    }
    ifl::CaseState s{};
    int st = readState(h, s);   // synchronises the stream (the field copies above too)
    if (st != IFL_OK) return st;
    ifl_state_header hd{};
    std::memcpy(hd.magic, IFL_STATE_MAGIC, 8);
    hd.version = IFL_STATE_VERSION;
    hd.dim = c.dim; hd.nx = c.nx; hd.ny = c.ny; hd.nz = c.nz;
    hd.last_converged = s.lastConverged;
    hd.iter = s.iter;
    const double nm[7] = {s.pNormMax, s.uNormMax, s.vNormMax, s.wNormMax, s.ufNormMax, s.vfNormMax, s.wfNormMax};
    for (int q = 0; q < 7; ++q) hd.norm_max[q] = nm[q];
    for (int q = 0; q < 5; ++q) hd.res_max[q] = s.resMax[q];
    hd.last_scaled = s.lastScaled;
    hd.payload_bytes = data.size() * 4;
    hd.checksum = fnv1a(reinterpret_cast<const unsigned char *>(data.data()), data.size() * 4);
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(IFL_ERR_IO, std::string("cannot open for writing: ") + path);
    const bool ok = std::fwrite(&hd, sizeof(hd), 1, fp) == 1
                    && std::fwrite(data.data(), 4, data.size(), fp) == data.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) return fail(IFL_ERR_IO, std::string("write failed: ") + path);
    return IFL_OK;
}

int ifl_load_state(ifl_solver *h, const char *path) {
    if (!h || !path) return fail(IFL_ERR_ARG, "null argument");
    const ifl_config &c = h->cfg;
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(IFL_ERR_IO, std::string("cannot open: ") + path);
    ifl_state_header hd{};
    std::vector<float> data(statePayloadFloats(c));
    const bool gotHeader = std::fread(&hd, sizeof(hd), 1, fp) == 1;
    int st = IFL_OK;
    if (!gotHeader || std::memcmp(hd.magic, IFL_STATE_MAGIC, 8) != 0)
        st = fail(IFL_ERR_IO, std::string("not a state file: ") + path);
    else if (hd.version != IFL_STATE_VERSION)
        st = fail(IFL_ERR_IO, "state file version " + std::to_string(hd.version) + ", expected "
                                  + std::to_string(IFL_STATE_VERSION));
    else if (hd.dim != c.dim || hd.nx != c.nx || hd.ny != c.ny || hd.nz != c.nz)
        st = fail(IFL_ERR_ARG, "state file grid " + std::to_string(hd.dim) + "D " + std::to_string(hd.nx) + "x"
                                   + std::to_string(hd.ny) + "x" + std::to_string(hd.nz) + " does not match the handle's");
    else if (hd.payload_bytes != data.size() * 4 || std::fread(data.data(), 4, data.size(), fp) != data.size()
             || std::fgetc(fp) != EOF)
        st = fail(IFL_ERR_IO, std::string("truncated or oversized state file: ") + path);
    else if (fnv1a(reinterpret_cast<const unsigned char *>(data.data()), data.size() * 4) != hd.checksum)
        st = fail(IFL_ERR_IO, std::string("state file checksum mismatch: ") + path);
    std::fclose(fp);
    if (st != IFL_OK) return st;
    ifl::CaseState s{};
    if ((st = readState(h, s)) != IFL_OK) return st;
    s.iter = hd.iter;
    s.lastConverged = hd.last_converged;
    double *nm[7] = {&s.pNormMax, &s.uNormMax, &s.vNormMax, &s.wNormMax, &s.ufNormMax, &s.vfNormMax, &s.wfNormMax};
    for (int q = 0; q < 7; ++q) *nm[q] = hd.norm_max[q];
    for (int q = 0; q < 5; ++q) s.resMax[q] = hd.res_max[q];
    s.lastScaled = hd.last_scaled;
    s.done = 0;
    size_t off = 0;
    for (int f : kStateFields) {
        const size_t n = fieldSize(c, f);
        IFL_HIP(hipMemcpyAsync(fieldPtr(h, f), data.data() + off, n * 4, hipMemcpyHostToDevice, h->stream));
        if (f == IFL_FIELD_U || f == IFL_FIELD_V || f == IFL_FIELD_W)   // the smoothers' partner buffers
            IFL_HIP(hipMemcpyAsync(h->P.alt[f - IFL_FIELD_U], data.data() + off, n * 4, hipMemcpyHostToDevice,
                                   h->stream));
        off += n;
    }
    IFL_HIP(hipMemcpyAsync(h->P.st, &s, sizeof(s), hipMemcpyHostToDevice, h->stream));
    IFL_HIP(hipStreamSynchronize(h->stream));
    h->iter = hd.iter;
    return IFL_OK;
}

}  // extern "C"
