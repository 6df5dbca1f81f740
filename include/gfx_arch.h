// gfx_arch.h -- compile-time limits of the AMD targets the assembly
// kernels are written for, keyed by a gfx code (950 = gfx950 / MI355X).
//
// The gfx950 counterpart of the reference's cuda_arch.h (smem_limits,
// fits(), needs_dynamic_smem_opt_in(): cuda_arch.h:37-76,124-129). On CDNA
// there is no static/dynamic LDS split to opt into and no carveout: a
// workgroup may allocate up to the whole LDS of its CU, so the launch
// planning reduces to "does the footprint fit" and "how many workgroups
// share a CU", plus the register budget a wave64 row-per-lane factorisation
// needs.
#ifndef GFX_ARCH_H
#define GFX_ARCH_H

#include <cstddef>

namespace gfx_arch {

inline constexpr std::size_t KiB = 1024;

struct limits {
    unsigned code;            // 950 for gfx950
    const char* name;         // offload-arch string
    unsigned wave_size;       // lanes per wavefront
    unsigned cus;             // compute units (MI355X: 256, in 8 XCDs)
    unsigned xcds;            // accelerator complex dies; blocks go round-robin
    unsigned simds_per_cu;    // SIMD units per CU
    unsigned lds_per_cu_kb;   // LDS per CU
    unsigned lds_per_wg_kb;   // largest LDS allocation of one workgroup
    unsigned vgprs_per_simd;  // 32-bit VGPRs per lane per SIMD (arch + acc)
    unsigned max_vgprs_wave;  // 32-bit VGPRs one wave may hold (arch + acc)

    constexpr std::size_t lds_per_wg() const noexcept { return lds_per_wg_kb * KiB; }
    constexpr bool fits(std::size_t bytes) const noexcept { return bytes <= lds_per_wg(); }
    // Never true on CDNA: kept so code written against the reference's
    // launch contract compiles unchanged.
    constexpr bool needs_dynamic_smem_opt_in(std::size_t) const noexcept { return false; }
    // Workgroups that fit one CU's LDS at this footprint.
    constexpr unsigned wgs_per_cu_by_lds(std::size_t bytes) const noexcept {
        return bytes ? static_cast<unsigned>((lds_per_cu_kb * KiB) / bytes) : 0u;
    }
    // Waves per SIMD at this VGPR count (granule 8).
    constexpr unsigned waves_per_simd_by_vgprs(unsigned vgprs) const noexcept {
        const unsigned g = (vgprs + 7u) / 8u * 8u;
        return g ? (vgprs_per_simd / g < 8u ? vgprs_per_simd / g : 8u) : 8u;
    }
};

inline constexpr limits table[] = {
    // code name      wave CUs XCD SIMD LDS/CU LDS/WG VGPR/SIMD VGPR/wave
    {950, "gfx950", 64, 256, 8, 4, 160, 160, 512, 512},
};

constexpr const limits* find(unsigned code) noexcept {
    for (const auto& e : table)
        if (e.code == code) return &e;
    return nullptr;
}

template <unsigned CODE>
inline constexpr limits limits_v = [] {
    static_assert(find(CODE) != nullptr, "not a supported gfx code");
    return *find(CODE);
}();

static_assert(limits_v<950>.wave_size == 64);
static_assert(limits_v<950>.fits(160 * KiB) && !limits_v<950>.fits(160 * KiB + 1));
static_assert(limits_v<950>.waves_per_simd_by_vgprs(128) == 4);

}  // namespace gfx_arch

#endif  // GFX_ARCH_H
