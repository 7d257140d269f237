// Storage of a 32-slot bank's half-precision inverse for the MFMA local
// solve: the 96 x 96 symmetric inverse (scaled by the bank's largest entry)
// as its 21 upper 16 x 16 tiles T(I, J), I <= J (diagonal tiles stored in
// full), each tile in MFMA lane order.
//
// A tile takes 512 bytes: 32 lanes x 8 halves, lane m holding row m % 16,
// columns 8 (m / 16) .. 8 (m / 16) + 7, i.e. the half index of element
// (r, c) is 8 ((c / 8) 16 + r) + c % 8.  That is the A-operand register
// layout of a 16 x 16 tile in one half of a wave for a 16x16x32 MFMA (lane l
// of the wave: row l % 16, k = 8 (l / 16) + 0..7), so one 16-byte load per
// lane fetches an operand.  Tiles are ordered so that the two tiles (I, 2P)
// and (I, 2P + 1) of every operand with I <= 2P lie next to each other
// (one 1 KB load per wave); the diagonal tiles (1,1), (3,3), (5,5) close the
// bank, since they are the upper half of an operand whose lower half is a
// transposed tile.
#pragma once

namespace masx
{
constexpr int kTileDim       = 16;
constexpr int kTileHalves    = kTileDim * kTileDim;        // 256
constexpr int kTilesPerBank  = 21;                          // 6 x 7 / 2
constexpr int kTileBankHalves = kTilesPerBank * kTileHalves;  // 5376 (10,752 bytes)
constexpr int kTilePairs     = 9;                           // operands with both tiles stored (I <= 2P)

// Position (0..20) of stored tile (I, J), I <= J < 6, in a bank.
__host__ __device__ constexpr int tile_pos(int I, int J)
{
    // pairs in operand order: (0,0)(0,1) (0,2)(0,3) (0,4)(0,5) (1,2)(1,3)
    // (1,4)(1,5) (2,2)(2,3) (2,4)(2,5) (3,4)(3,5) (4,4)(4,5), then (1,1)
    // (3,3) (5,5)
    return I == 0   ? J
           : I == 1 ? (J == 1 ? 18 : 4 + J)
           : I == 2 ? 8 + J
           : I == 3 ? (J == 3 ? 19 : 10 + J)
           : I == 4 ? 12 + J
                    : 20;
}

// Load slot (1 KB, two tiles) of the operand of block row I and column pair
// P when I <= 2P: the slot whose first tile is (I, 2P).
__host__ __device__ constexpr int tile_pair_slot(int I, int P) { return tile_pos(I, 2 * P) / 2; }

// Half index, within a bank, of the stored element (r, c) of the 96 x 96
// inverse with r <= c (the upper triangle; a diagonal tile's lower elements
// are stored too, at their own (r, c)).
__host__ __device__ constexpr int tile_half_index(int r, int c)
{
    return tile_pos(r / kTileDim, c / kTileDim) * kTileHalves +
           8 * (((c % kTileDim) / 8) * kTileDim + r % kTileDim) + (c % 8);
}
}  // namespace masx
