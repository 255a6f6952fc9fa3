// hpcs_tuning.hpp -- tuning tables of libhpcsort.so: one row per line,
//   <dtype> <max_n> <algo> <shape> <rank> [<digit_bits>]
// or, for the batched segmented sort,
//   segs <max_n> <small_max> <rank> [<shape>]
// Plain C++ (no device code), so the parser is tested on the host with
// sanitizers.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "hpcs_shapes.hpp"

namespace hpcs {

// One row of a tuning table: for sorts of up to max_n keys of key_bytes
// (with or without u32 values), the pass algorithm, tile shape, ranking and
// digit width.  A `segs` row (key_bytes 0): for batched segmented sorts of
// up to max_n keys in all, the largest segment sorted in LDS (seg_small), the
// ranking of the large segments' scatter (rank) and their tile shape (shape,
// a row of TILE_SHAPES with seg_shape_ok; 0 if the row omits it).
struct TuneEntry {
    int key_bytes;
    int vals;
    int64_t max_n;
    int algo, shape, rank;
    int bits;
    int seg_small;
};

// The sizes of segment that one workgroup may sort in LDS (hpcs_segments.hpp:
// SEG_SMALL, SEG_MEDIUM).
inline bool seg_small_valid(long long v) { return v == 4096 || v == 16384; }

// Parse a tuning table: one row per line, `#` starts a comment,
//   <dtype> <max_n> <algo> <shape> <rank> [<digit_bits>]
// dtype is u32, f32, i32 (4-byte keys), u64, f64, i64 (8-byte keys), pairs32
// or pairs64 (keys with u32 values); algo 0 onesweep / 1 reduce-then-scan /
// 3 one workgroup in LDS (arrays of at most 16384 4-byte or 8192 8-byte keys;
// larger arrays fall back to the choice by size) / 4 one MSD split, then one
// workgroup per bucket (keys without values, at most 2^22 4-byte or 2^21
// 8-byte keys; larger arrays fall back to the choice by size) / 5 runs of
// 4096 4-byte or 2048 8-byte keys sorted in LDS, then stable merge passes
// (arrays of at most 2^20 keys; larger ones fall back to the choice by size)
// / 6 the MSD-first form of 4-byte keys without values (top byte, then byte
// 2, then each (byte 3, byte 2) bucket in LDS; unsigned keys only, float and
// signed keys of the row take algo 1 with its shape and ranking);
// shape: a row of TILE_SHAPES (hpcs_shapes.hpp) that serves the dtype; rank
// 0 ballot / 1 LDS-OR / 2 and 3 LDS-OR with every second / fourth item by
// ballot (0 only for ballot-only shapes); digit_bits 8 (default) or 11.  11-bit digits exist in one form
// only: onesweep passes, 256x16 tiles, ballot ranking, keys without values,
// so a row with digit_bits 11 reads `<dtype> <max_n> 0 0 0 11`.
inline bool parse_tuning(const std::string &text, std::vector<TuneEntry> &out, std::string &err)
{
    out.clear();
    size_t pos = 0;
    int line_no = 0;
    while (pos <= text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        ++line_no;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        char dt[16] = {0};
        long long max_n = 0;
        int algo = 0, shape = 0, rank = 0, bits = 8, consumed = 0, consumed6 = 0;
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        if (std::sscanf(line.c_str(), " %15s", dt) == 1 && std::string(dt) == "segs") {
            long long small_max = 0;
            int consumed5 = 0;
            const int got = std::sscanf(line.c_str(), " %15s %lld %lld %d %n%d %n", dt, &max_n, &small_max, &rank,
                                        &consumed, &shape, &consumed5);
            if (got == 5) consumed = consumed5;
            if (got < 4 || line.find_first_not_of(" \t\r", size_t(consumed)) != std::string::npos) {
                err = "line " + std::to_string(line_no) + ": expected segs <max_n> <small_max> <rank> [<shape>]";
                return false;
            }
            if (got < 5) shape = 0;
            if (max_n < 1 || !seg_small_valid(small_max) || rank < 0 || rank > 3 || !seg_shape_ok(shape)) {
                err = "line " + std::to_string(line_no) + ": value out of range";
                return false;
            }
            TuneEntry e{};
            e.key_bytes = 0;
            e.max_n = max_n;
            e.rank = rank;
            e.shape = shape;
            e.bits = 8;
            e.seg_small = int(small_max);
            out.push_back(e);
            continue;
        }
        const int got = std::sscanf(line.c_str(), " %15s %lld %d %d %d %n%d %n", dt, &max_n, &algo, &shape, &rank,
                                    &consumed, &bits, &consumed6);
        if (got == 6) consumed = consumed6;
        if (got < 5 || line.find_first_not_of(" \t\r", size_t(consumed)) != std::string::npos) {
            err = "line " + std::to_string(line_no) + ": expected <dtype> <max_n> <algo> <shape> <rank> [<digit_bits>]";
            return false;
        }
        const std::string d(dt);
        TuneEntry e{};
        if (d == "u32" || d == "f32" || d == "i32") e.key_bytes = 4, e.vals = 0;
        else if (d == "u64" || d == "f64" || d == "i64") e.key_bytes = 8, e.vals = 0;
        else if (d == "pairs32") e.key_bytes = 4, e.vals = 1;
        else if (d == "pairs64") e.key_bytes = 8, e.vals = 1;
        else {
            err = "line " + std::to_string(line_no) + ": unknown dtype " + d;
            return false;
        }
        if (max_n < 1 || (algo != 0 && algo != 1 && algo != 3 && algo != 4 && algo != 5 && algo != 6) || rank < 0 ||
            rank > 3 || !shape_serves(shape, e.key_bytes, e.vals != 0) ||
            (rank != 0 && TILE_SHAPES[shape].ballot_only) || (bits != 8 && bits != 11) ||
            (bits == 11 && (algo != 0 || shape != 0 || rank != 0 || e.vals)) || (algo == 4 && e.vals) ||
            (algo == 6 && (e.vals || e.key_bytes != 4 || bits != 8))) {
            err = "line " + std::to_string(line_no) + ": value out of range";
            return false;
        }
        e.max_n = max_n;
        e.algo = algo;
        e.shape = shape;
        e.rank = rank;
        e.bits = bits;
        out.push_back(e);
    }
    return true;
}

} // namespace hpcs
