// Binary snapshot files (planned in the reference, plan.md:160-163; the layout is this build's,
// described with tpm_snapshot_header in tpm.h). Host code only.
#include <cstdio>
#include <cstring>

#include "tpm.h"

static_assert(sizeof(tpm_snapshot_header) == 128, "the snapshot header is 128 bytes");

namespace {

constexpr char kMagic[8] = {'T', 'P', 'M', 'S', 'N', 'A', 'P', '1'};
constexpr uint32_t kVersion = 1;

struct File {
    std::FILE *f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

bool host_little_endian() {
    const uint32_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

}  // namespace

extern "C" {

int tpm_snapshot_write(const char *path, const tpm_snapshot_header *header, const double *pos, const double *mom) {
    if (!path || !header || !host_little_endian()) return TPM_E_INVALID;
    if (header->n > 0 && (!pos || !mom)) return TPM_E_INVALID;
    tpm_snapshot_header h = *header;
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    h.header_bytes = (uint32_t)sizeof(tpm_snapshot_header);
    h.reserved = 0;
    File out;
    out.f = std::fopen(path, "wb");
    if (!out.f) return TPM_E_IO;
    const size_t m = (size_t)(3 * h.n);
    if (std::fwrite(&h, sizeof(h), 1, out.f) != 1) return TPM_E_IO;
    if (m && std::fwrite(pos, sizeof(double), m, out.f) != m) return TPM_E_IO;
    if (m && std::fwrite(mom, sizeof(double), m, out.f) != m) return TPM_E_IO;
    const int rc = std::fclose(out.f);
    out.f = nullptr;
    return rc == 0 ? TPM_OK : TPM_E_IO;
}

int tpm_snapshot_read(const char *path, tpm_snapshot_header *header, double *pos, double *mom, size_t capacity) {
    if (!path || !header || !host_little_endian()) return TPM_E_INVALID;
    File in;
    in.f = std::fopen(path, "rb");
    if (!in.f) return TPM_E_IO;
    tpm_snapshot_header h;
    if (std::fread(&h, sizeof(h), 1, in.f) != 1) return TPM_E_IO;
    if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0 || h.version != kVersion ||
        h.header_bytes != sizeof(tpm_snapshot_header))
        return TPM_E_IO;
    if ((pos || mom) && capacity < h.n) return TPM_E_INVALID;
    const size_t m = (size_t)(3 * h.n);
    if (pos) {
        if (m && std::fread(pos, sizeof(double), m, in.f) != m) return TPM_E_IO;
    } else if (m && std::fseek(in.f, (long)(m * sizeof(double)), SEEK_CUR) != 0) {
        return TPM_E_IO;
    }
    if (mom && m && std::fread(mom, sizeof(double), m, in.f) != m) return TPM_E_IO;
    *header = h;
    return TPM_OK;
}

}  // extern "C"
