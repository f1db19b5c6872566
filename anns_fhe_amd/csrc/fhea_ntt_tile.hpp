// fhea_ntt_tile.hpp -- tile geometry shared by the FP64 and the integer transforms.
//
// A workgroup of 256 threads owns a 4096-word tile of one limb; each thread keeps
// 16 tile elements in registers, the elements whose row index differs only in a
// 4-bit window, and runs up to four radix-2 stages on them per round; rounds
// exchange through LDS.  Round<> gives a round's window and stages at compile
// time; find_tile walks a persistent grid over the tiles of a batch.
#pragma once
#include "fhea_kernels.hpp"

namespace fhea {

// The thread index as a value the compiler cannot hoist: addresses derived from it
// inside a loop are recomputed there (a few integer instructions) instead of being
// kept live across the loop (dozens of registers for the exchanges' LDS addresses).
__device__ __forceinline__ u32 opaque_tid(u32 v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Source index of word i of sigma_g(a) in the bit-reversed evaluation order:
//   map[i] = br(((g (2 br(i) + 1)) mod 2N - 1) / 2)
// (the image of a(X) -> a(X^g); automorphism.cpp:3-12 reads a[map[i]]).
__device__ __forceinline__ u32 galois_src(u32 i, u64 g, u32 logN) {
    const u32 br = __brev(i) >> (32 - logN);
    const u64 e = (g * (2ull * br + 1)) & ((2ull << logN) - 1);
    const u32 k = (u32)((e - 1) >> 1);
    return __brev(k) >> (32 - logN);
}

// LDS index of tile element e, padded by one word per 16 (bank spread)
__device__ __forceinline__ u32 lds_pad(u32 e) { return e + (e >> 4); }

template <bool INV, int LOGR, int RD>
struct Round {
    static constexpr int NR = (LOGR + 3) / 4;
    static constexpr int NS = (RD + 1 < NR) ? 4 : (LOGR - 4 * RD);  // stages in this round
    static constexpr int TOP = LOGR - 1 - 4 * RD;
    // window [P, P+4) of row-index bits; BIT0 = bit of the round's first stage
    static constexpr int P = INV ? ((4 * RD + 4 <= LOGR) ? 4 * RD : LOGR - 4) : (TOP >= 3 ? TOP - 3 : 0);
    static constexpr int BIT0 = INV ? 4 * RD : TOP;
};

// global word index of tile element e in a limb viewed as [blocks][2^logR][2^logS],
// the tile covering 2^logC adjacent columns from c0 and blocks from b0
__device__ __forceinline__ u32 tile_gidx(u32 e, u32 logR, u32 logS, u32 logC, u32 b0, u32 c0) {
    const u32 c = e & ((1u << logC) - 1), r = (e >> logC) & ((1u << logR) - 1), ob = e >> (logC + logR);
    return ((b0 + ob) << (logR + logS)) + (r << logS) + c0 + c;
}

// Work item = one 4096-word tile (tiles_per_limb per limb slot, n_tiles in all).
// The kernels are persistent: workgroup w takes tiles w, w + G, w + 2G, ...
// Tiles of skipped slots (prime map 0xFFFFFFFF) are passed over, and so are the
// tiles of primes outside pmask when one is given (bit p of pmask[p / 64]: a
// chain whose transforms run partly on the FP64 and partly on the integer kernel).
struct TileId {
    u32 tile, poly, slot, pidx, tI;
};

// Tiles are numbered (limb slot, polynomial, tile-in-limb) with the polynomial
// varying fastest after the tile: the same prime's limbs of a whole batch are
// in flight together and share its twiddle table in L2.
__device__ __forceinline__ TileId find_tile(u32 t, u32 n_tiles, u32 tiles_per_limb, u32 npolys,
                                            const u32 *prime_map, const u32 *slots, const u64 *pmask) {
    TileId id{n_tiles, 0, 0, 0, 0};
    for (; t < n_tiles; t += gridDim.x) {
        const u32 yl = t / tiles_per_limb;
        const u32 k = yl / npolys;
        const u32 slot = slots ? slots[k] : k;
        const u32 p = prime_map[slot];
        if (p != 0xFFFFFFFFu && (!pmask || ((pmask[p >> 6] >> (p & 63)) & 1))) {
            id.tile = t;
            id.poly = yl % npolys;
            id.slot = slot;
            id.pidx = p;
            id.tI = t % tiles_per_limb;
            return id;
        }
    }
    return id;
}

// Unit-major order for prologues that read one word per position shared by every
// slot (the rescale lift v): unit u = (polynomial, tile-in-limb), and the nslots
// tiles of a unit are numbered 8 apart inside groups of 8 units,
//   t = (u / 8) * 8 * nslots + slot * 8 + u % 8,
// so on a persistent grid of a multiple of 8 workgroups (dispatched round-robin
// over the 8 XCDs) all slots of a unit run on one XCD and read its v words from
// that XCD's L2.  n_tiles = ceil(units / 8) * 8 * nslots; indices of units past
// the end are skipped.
__device__ __forceinline__ TileId find_tile_um(u32 t, u32 n_tiles, u32 tiles_per_limb, u32 npolys, u32 nslots,
                                               const u32 *prime_map, const u32 *slots, const u64 *pmask) {
    TileId id{n_tiles, 0, 0, 0, 0};
    const u32 units = npolys * tiles_per_limb;
    for (; t < n_tiles; t += gridDim.x) {
        const u32 rem = t % (8 * nslots);
        const u32 k = rem >> 3, u = (t / (8 * nslots)) * 8 + (rem & 7);
        if (u >= units) continue;
        const u32 slot = slots ? slots[k] : k;
        const u32 p = prime_map[slot];
        if (p != 0xFFFFFFFFu && (!pmask || ((pmask[p >> 6] >> (p & 63)) & 1))) {
            id.tile = t;
            id.poly = u / tiles_per_limb;
            id.slot = slot;
            id.pidx = p;
            id.tI = u % tiles_per_limb;
            return id;
        }
    }
    return id;
}

}  // namespace fhea
