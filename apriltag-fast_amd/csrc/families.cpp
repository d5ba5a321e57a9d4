// The built-in tag families (host code): the packed tables of
// families/families.bin, embedded in the library at build time, turned
// into at_family structs on first use. Mirrors the built-ins of the
// reference's native API (fast/families.c:37-78): the same names, order
// and field values.
#include <cstring>
#include <mutex>
#include <vector>

#include "atg.h"

#ifndef ATG_FAMILIES_BIN
#error "build with -DATG_FAMILIES_BIN=\"<absolute path of families.bin>\""
#endif

__asm__(".section .rodata\n"
        ".balign 16\n"
        ".global atg_family_blob\n"
        "atg_family_blob:\n"
        ".incbin \"" ATG_FAMILIES_BIN "\"\n"
        ".global atg_family_blob_end\n"
        "atg_family_blob_end:\n"
        ".byte 0\n"
        ".previous\n");

extern "C" const unsigned char atg_family_blob[];
extern "C" const unsigned char atg_family_blob_end[];

namespace {

const char *kNames[9] = {"tag16h5",        "tag25h9",        "tag36h10",
                         "tag36h11",       "tagCircle21h7",  "tagCircle49h12",
                         "tagCustom48h12", "tagStandard41h12", "tagStandard52h13"};

struct Builtins {
    at_family fam[9];
    std::vector<uint64_t> codes[9];
    bool ok = false;
};

Builtins &builtins() {
    static Builtins b;
    static std::once_flag once;
    std::call_once(once, [] {
        const unsigned char *p = atg_family_blob;
        size_t len = (size_t)(atg_family_blob_end - atg_family_blob);
        auto u32 = [&](size_t at) { uint32_t v; memcpy(&v, p + at, 4); return v; };
        if (len < 12 || memcmp(p, "ATGF", 4) != 0 || u32(4) != 1 || u32(8) != 9) return;
        size_t at = 12;
        for (int f = 0; f < 9; f++) {
            at_family &d = b.fam[f];
            memset(&d, 0, sizeof d);
            memcpy(d.name, p + at, 31);
            at += 32;
            uint32_t ncodes = u32(at), nbits = u32(at + 4), h = u32(at + 8);
            int32_t wab = (int32_t)u32(at + 12), tw = (int32_t)u32(at + 16);
            uint32_t rev = u32(at + 20);
            at += 24;
            if (nbits > 64 || strcmp(d.name, kNames[f]) != 0) return;
            d.nbits = (int32_t)nbits;
            d.h = (int32_t)h;
            d.width_at_border = wab;
            d.total_width = tw;
            d.reversed_border = rev != 0;
            for (uint32_t i = 0; i < nbits; i++) d.bit_x[i] = (int32_t)u32(at + 4 * i);
            at += 4 * nbits;
            for (uint32_t i = 0; i < nbits; i++) d.bit_y[i] = (int32_t)u32(at + 4 * i);
            at += 4 * nbits;
            b.codes[f].resize(ncodes);
            memcpy(b.codes[f].data(), p + at, 8 * (size_t)ncodes);
            at += 8 * (size_t)ncodes;
            d.codes = b.codes[f].data();
            d.ncodes = ncodes;
        }
        b.ok = at <= len;
    });
    return b;
}

}  // namespace

extern "C" const at_family *atg_builtin_family(int32_t index) {
    Builtins &b = builtins();
    if (!b.ok || index < 0 || index >= 9) return nullptr;
    return &b.fam[index];
}

extern "C" const at_family *at_family_by_name(const char *name) {
    for (int i = 0; i < 9; i++)
        if (strcmp(kNames[i], name) == 0) return atg_builtin_family(i);
    return nullptr;
}

extern "C" const at_family *at_tag16h5(void) { return atg_builtin_family(0); }
extern "C" const at_family *at_tag25h9(void) { return atg_builtin_family(1); }
extern "C" const at_family *at_tag36h10(void) { return atg_builtin_family(2); }
extern "C" const at_family *at_tag36h11(void) { return atg_builtin_family(3); }
extern "C" const at_family *at_tagCircle21h7(void) { return atg_builtin_family(4); }
extern "C" const at_family *at_tagCircle49h12(void) { return atg_builtin_family(5); }
extern "C" const at_family *at_tagCustom48h12(void) { return atg_builtin_family(6); }
extern "C" const at_family *at_tagStandard41h12(void) { return atg_builtin_family(7); }
extern "C" const at_family *at_tagStandard52h13(void) { return atg_builtin_family(8); }
