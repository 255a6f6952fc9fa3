// hpcs_shapes.hpp -- the tile shapes of the radix sort passes, in one table.
//
// A tile is the keys one workgroup takes in a pass: BLOCK threads x ITEMS
// keys per thread.  Every kernel of a pass (upsweep, tile counts, scatter /
// onesweep) is instantiated from this table through one dispatch
// (hpcs_api.hip, with_shape), so the counting and scattering kernels of a
// sort always agree on the tile, and a row without kernels cannot exist: the
// dispatch instantiates every row it may select.  Plain C++ (no device code):
// the tuning-table parser validates shapes against it on the host.
#pragma once

namespace hpcs {

// Which sorts a shape serves: the staging area of a tile (keys, and u32
// values where the sort has them) must fit the shape's LDS budget.
constexpr unsigned SERVE_K4 = 1u;   // 4-byte keys without values
constexpr unsigned SERVE_K4V = 2u;  // 4-byte keys with u32 values
constexpr unsigned SERVE_K8 = 4u;   // 8-byte keys without values
constexpr unsigned SERVE_K8V = 8u;  // 8-byte keys with u32 values
constexpr unsigned SERVE_ALL = SERVE_K4 | SERVE_K4V | SERVE_K8 | SERVE_K8V;

struct TileShape {
    int block, items;  // threads x keys per thread
    unsigned serves;   // SERVE_* mask
    bool ballot_only;  // the LDS-OR match masks do not fit beside the staging area
};

// 0-4: 4096 / 8192 / 16384 / 2048-key tiles; 5 and 6: 1024-thread
// workgroups whose 64 KiB staging area and 16-bit per-wave counters leave
// room for two workgroups per CU (ballot ranking only); 7-9: 256 threads x
// 12 / 20 / 24 keys.
constexpr TileShape TILE_SHAPES[] = {
    {256, 16, SERVE_ALL, false},              // 0
    {512, 16, SERVE_K4, false},               // 1
    {256, 32, SERVE_K4, false},               // 2
    {512, 32, SERVE_K4, false},               // 3
    {256, 8, SERVE_ALL, false},               // 4
    {1024, 16, SERVE_K4, true},               // 5: 16384 4-byte keys
    {1024, 8, SERVE_K4V | SERVE_K8, true},    // 6: 8192 keys (4-byte with values, or 8-byte)
    {256, 12, SERVE_K4, false},               // 7
    {256, 20, SERVE_K4, false},               // 8
    {256, 24, SERVE_K4, false},               // 9
};
constexpr int NUM_TILE_SHAPES = int(sizeof(TILE_SHAPES) / sizeof(TILE_SHAPES[0]));

constexpr unsigned serve_bit(int key_bytes, bool vals)
{
    return key_bytes == 4 ? (vals ? SERVE_K4V : SERVE_K4) : (vals ? SERVE_K8V : SERVE_K8);
}

// Shape s can sort keys of key_bytes (4 or 8), with or without u32 values.
constexpr bool shape_serves(int s, int key_bytes, bool vals)
{
    return s >= 0 && s < NUM_TILE_SHAPES && (TILE_SHAPES[s].serves & serve_bit(key_bytes, vals)) != 0;
}

constexpr int shape_tile_keys(int s) { return TILE_SHAPES[s].block * TILE_SHAPES[s].items; }

// Shapes of the batched segments' large-segment tiles: the 256- and
// 512-thread rows for 4-byte keys without values (LDS-OR or ballot ranking).
constexpr bool seg_shape_ok(int s) { return shape_serves(s, 4, false) && TILE_SHAPES[s].block <= 512; }

static_assert(shape_serves(0, 4, false) && shape_serves(0, 4, true) && shape_serves(0, 8, false) &&
                  shape_serves(0, 8, true),
              "shape 0 is the default of every sort");

} // namespace hpcs
