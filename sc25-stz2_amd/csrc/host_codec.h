// host_codec.h -- host-side pieces of the STZ path: Huffman codebook
// construction, decode tables, base-codec pass schedule, ROI planner.
#pragma once
#include <stdint.h>

#include <vector>

#include "kernels.h"
#include "layout.h"

namespace stz {

// ---- Huffman ----------------------------------------------------------------
struct Codebook {
    // used symbols in ascending symbol order
    std::vector<uint16_t> sym;
    std::vector<uint8_t> len;
    std::vector<uint32_t> code;
    uint64_t total_bits = 0;
};

// freq pairs (symbol, count) in ascending symbol order -> canonical, length-limited code
void build_codebook(const uint64_t* sym_freq, uint32_t n, int lmax, Codebook& cb);
// lengths for a dense histogram (for the stage entry point)
int huff_lengths_dense(const uint64_t* freq, int nsym, int lmax, uint8_t* len);
// canonical decode description from (symbol, length) pairs in symbol order
// (the table itself is filled on the device from it)
void build_dechead(const uint16_t* sym, const uint8_t* len, uint32_t n, DecHead& h, std::vector<uint16_t>& sorted);

// ---- base codec schedule ---------------------------------------------------
int interp_levels(int bx, int by, int bz);
double level_div(double alpha, double beta, int l);
// every pass of the base codec in stream order; stop_level > 0 drops the
// levels finer than stop_level
std::vector<InterpPass> make_passes(int bx, int by, int bz, double eb, int interp, int order,
                                    double alpha, double beta, int stop_level);

// ---- ROI / slice planner (F5:1578-1830) -----------------------------------
struct Box { int x0, x1, y0, y1, z0, z1; };
struct BlockPlan { bool valid = false; Box r{}; std::vector<int64_t> chunks; };
struct Plan {
    Box q{};
    bool is_slice = false;
    BlockPlan high[7], low[7];
    std::vector<Box> regions;  // mid-grid regions the high level reads
};
// nx..: caller dims (queries are normalised against them); px..: padded dims
void make_plan(Plan& pl, int nx, int ny, int nz, int px, int py, int pz, int chunk_high,
               int chunk_low, int ax, int ay, int az, int bx, int by, int bz);

}  // namespace stz
