// Host-only check of the work-image re-cutter (flashx_amd/csrc/retile.cpp):
// reads an image and its offsets, validates it as an upload does, re-cuts it
// into h2-row tiles and writes the new image and offsets.  No GPU call is
// made.  Usage: retile_check in.img in.offs h2 out.img out.offs
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../flashx_amd/csrc/fmx_internal.h"

static bool read_all(const char *path, std::vector<uint8_t> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize((size_t)n);
    bool ok = n == 0 || fread(buf.data(), (size_t)n, 1, f) == 1;
    fclose(f);
    return ok;
}

int main(int argc, char **argv)
{
    if (argc != 6) return 2;
    fmx_spm s;
    std::vector<uint8_t> offb;
    if (!read_all(argv[1], s.h_image) || !read_all(argv[2], offb)) return 3;
    s.offs.resize(offb.size() / 8);
    memcpy(s.offs.data(), offb.data(), offb.size());
    if (fmx::analyse_image(nullptr, &s) != FMX_OK) return 4;
    std::vector<uint8_t> img;
    std::vector<int64_t> offs;
    fmx::retile_image(&s, (uint32_t)atoi(argv[3]), img, offs);
    fmx_spm t;
    t.h_image = img;
    t.offs = offs;
    if (fmx::analyse_image(nullptr, &t) != FMX_OK || t.nnz != s.nnz) return 5;
    FILE *f = fopen(argv[4], "wb");
    FILE *g = fopen(argv[5], "wb");
    if (!f || !g) return 6;
    fwrite(img.data(), img.size(), 1, f);
    fwrite(offs.data(), offs.size() * 8, 1, g);
    fclose(f);
    fclose(g);
    printf("OK %zu bytes, %zu block rows\n", img.size(), offs.size() - 1);
    return 0;
}
