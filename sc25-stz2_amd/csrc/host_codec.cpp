// host_codec.cpp -- Huffman codebook, decode tables, base-codec schedule and
// the ROI/slice planner.
//
// Huffman: the reference calls SZ3-ncp's HuffmanEncoder (F5:1271, F5:1402),
// whose tree construction and tie-breaking are not available.  This build's
// code: two-queue Huffman over leaves sorted by (frequency, symbol) with ties
// going to the leaf queue; lengths above the limit are repaired with the
// Annex K.3 procedure of ITU-T T.81 on the per-length counts and handed back
// in (old length, symbol) order; canonical codes are assigned DEFLATE style
// (RFC 1951 sec. 3.2.2) over (length, symbol) order.
#include "host_codec.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>

namespace stz {

namespace {

struct Leaf { uint64_t f; uint32_t sym; };

// lengths for leaves (symbol, freq) given in ascending symbol order
void lengths_for(const std::vector<Leaf>& in, int lmax, std::vector<uint8_t>& out_len)
{
    const int n = (int)in.size();
    out_len.assign(n, 0);
    if (n == 0) return;
    if (n == 1) { out_len[0] = 1; return; }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        if (in[a].f != in[b].f) return in[a].f < in[b].f;
        return in[a].sym < in[b].sym;
    });
    std::vector<uint64_t> w(2 * n);
    std::vector<int> parent(2 * n, -1);
    for (int i = 0; i < n; ++i) w[i] = in[idx[i]].f;
    int li = 0, ii = n, next = n;
    for (int k = 0; k < n - 1; ++k) {
        int pick[2];
        for (int t = 0; t < 2; ++t) {
            if (li < n && (ii >= next || w[li] <= w[ii])) pick[t] = li++;
            else pick[t] = ii++;
        }
        w[next] = w[pick[0]] + w[pick[1]];
        parent[pick[0]] = next;
        parent[pick[1]] = next;
        next++;
    }
    std::vector<int> depth(2 * n, 0);
    for (int k = next - 2; k >= n; --k) depth[k] = depth[parent[k]] + 1;
    int maxlen = 0;
    for (int i = 0; i < n; ++i) {
        depth[i] = depth[parent[i]] + 1;
        maxlen = std::max(maxlen, depth[i]);
    }
    if (maxlen <= lmax) {
        for (int i = 0; i < n; ++i) out_len[idx[i]] = (uint8_t)depth[i];
        return;
    }
    // T.81 Annex K.3 on the counts per length
    std::vector<long> bits(maxlen + 2, 0);
    for (int i = 0; i < n; ++i) bits[depth[i]]++;
    for (int i = maxlen; i > lmax; --i) {
        while (bits[i] > 0) {
            int j = i - 2;
            while (bits[j] == 0) j--;
            bits[i] -= 2;
            bits[i - 1] += 1;
            bits[j + 1] += 2;
            bits[j] -= 1;
        }
    }
    // (old depth, symbol) order; position i in `in` is symbol order already
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::vector<int> dep_of(n);
    for (int i = 0; i < n; ++i) dep_of[idx[i]] = depth[i];
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return dep_of[a] < dep_of[b]; });
    int k = 0;
    for (int l = 1; l <= lmax; ++l)
        for (long c = 0; c < bits[l]; ++c) out_len[ord[k++]] = (uint8_t)l;
}

void canonical(const std::vector<uint8_t>& len, const std::vector<uint16_t>& sym, std::vector<uint32_t>& code)
{
    const int n = (int)len.size();
    code.assign(n, 0);
    uint32_t count[kLmax + 2] = {0};
    for (int i = 0; i < n; ++i) count[len[i]]++;
    uint32_t next_code[kLmax + 2] = {0};
    uint32_t c = 0;
    for (int l = 1; l <= kLmax; ++l) {
        c = (l == 1) ? 0u : (c + count[l - 1]) << 1;
        next_code[l] = c;
    }
    // symbols are in ascending order, so per-length assignment is in (len, sym) order
    for (int i = 0; i < n; ++i) code[i] = next_code[len[i]]++;
    (void)sym;
}

}  // namespace

void build_codebook(const uint64_t* sym_freq, uint32_t n, int lmax, Codebook& cb)
{
    std::vector<Leaf> leaves(n);
    cb.sym.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        leaves[i].sym = (uint32_t)(sym_freq[i] >> 32);
        leaves[i].f = sym_freq[i] & 0xFFFFFFFFull;
        cb.sym[i] = (uint16_t)leaves[i].sym;
    }
    lengths_for(leaves, lmax, cb.len);
    canonical(cb.len, cb.sym, cb.code);
    cb.total_bits = 0;
    for (uint32_t i = 0; i < n; ++i) cb.total_bits += leaves[i].f * cb.len[i];
}

int huff_lengths_dense(const uint64_t* freq, int nsym, int lmax, uint8_t* len)
{
    std::vector<Leaf> leaves;
    for (int s = 0; s < nsym; ++s)
        if (freq[s]) leaves.push_back({freq[s], (uint32_t)s});
    std::vector<uint8_t> l;
    lengths_for(leaves, lmax, l);
    std::memset(len, 0, (size_t)nsym);
    for (size_t i = 0; i < leaves.size(); ++i) len[leaves[i].sym] = l[i];
    return (int)leaves.size();
}

void build_dechead(const uint16_t* sym, const uint8_t* len, uint32_t n, DecHead& h, std::vector<uint16_t>& sorted)
{
    std::memset(&h, 0, sizeof(h));
    uint32_t count[kLmax + 2] = {0};
    for (uint32_t i = 0; i < n; ++i) count[len[i]]++;
    uint32_t c = 0, idx = 0;
    uint32_t fill[kLmax + 2] = {0};
    for (int l = 1; l <= kLmax; ++l) {
        c = (l == 1) ? 0u : (c + count[l - 1]) << 1;
        h.first_code[l] = c;
        h.count[l] = count[l];
        h.first_index[l] = idx;
        fill[l] = idx;
        idx += count[l];
    }
    // symbols of one length get consecutive codes in symbol order
    sorted.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) sorted[fill[len[i]]++] = sym[i];
}

int interp_levels(int bx, int by, int bz)
{
    const int m = std::max(bx, std::max(by, bz));
    int L = 0;
    while ((1 << L) < m) L++;
    return L;
}

double level_div(double alpha, double beta, int l)
{
    const double a = std::pow(alpha, (double)(l - 1));
    return a < beta ? a : beta;
}

std::vector<InterpPass> make_passes(int bx, int by, int bz, double eb, int interp, int order, double alpha,
                                    double beta, int stop_level)
{
    static const int tab[6][3] = {{2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 0, 2}, {0, 2, 1}, {0, 1, 2}};
    const int o = (order >= 0 && order < 6) ? order : 0;
    const int* ord = tab[o];
    const int dims[3] = {bx, by, bz};
    const int L = interp_levels(bx, by, bz);
    std::vector<InterpPass> out;
    uint64_t pos = 0;
    {
        InterpPass p{};
        p.axis = -1;
        p.cnt[0] = p.cnt[1] = p.cnt[2] = 1;
        p.step[0] = p.step[1] = p.step[2] = 1;
        p.bx = bx; p.by = by; p.bz = bz;
        p.interp = interp;
        p.eb = eb / level_div(alpha, beta, L + 1);
        p.recip = 1.0 / p.eb;
        p.pos = pos;
        p.n = 1;
        out.push_back(p);
        pos += 1;
    }
    for (int l = L; l >= 1; --l) {
        if (l < stop_level) break;
        const int s = 1 << (l - 1);
        const double ebl = eb / level_div(alpha, beta, l);
        for (int pi = 0; pi < 3; ++pi) {
            InterpPass p{};
            for (int q = 0; q < 3; ++q) {
                const int rank = (ord[0] == q) ? 0 : (ord[1] == q) ? 1 : 2;
                if (rank == pi) { p.start[q] = s; p.step[q] = 2 * s; }
                else if (rank < pi) { p.start[q] = 0; p.step[q] = s; }
                else { p.start[q] = 0; p.step[q] = 2 * s; }
                p.cnt[q] = dims[q] > p.start[q] ? (dims[q] - 1 - p.start[q]) / p.step[q] + 1 : 0;
            }
            p.axis = ord[pi];
            p.s = s;
            p.interp = interp;
            p.bx = bx; p.by = by; p.bz = bz;
            p.eb = ebl;
            p.recip = 1.0 / ebl;
            p.pos = pos;
            p.n = (uint64_t)p.cnt[0] * p.cnt[1] * p.cnt[2];
            pos += p.n;
            out.push_back(p);
        }
    }
    return out;
}

// ---- planner: restatement of F5:1578-1830 ----------------------------------
namespace {
struct AxisRange { int b, e; };

AxisRange normalize_axis(int a, int b, int dim, bool& is_slice)  // F5:1622-1635
{
    if (a == b) {
        is_slice = true;
        const int v = std::max(0, std::min(a, dim - 1));
        return {v, v};
    }
    int lo = std::max(0, std::min(a, b));
    int hi = std::min(dim - 1, std::max(a, b));
    if (hi < lo) hi = lo;
    return {lo, hi};
}

AxisRange range_for_parity(AxisRange f, int parity)  // F5:1646-1660
{
    int first = f.b;
    if ((first & 1) != parity) first++;
    if (first > f.e) return {1, 0};
    int last = f.e;
    if ((last & 1) != parity) last--;
    return {first / 2, last / 2};
}

// F5:1662-1687, with a partial last chunk per axis
void chunks_for_box(int dx, int dy, int dz, int C, AxisRange x, AxisRange y, AxisRange z, std::vector<int64_t>& out)
{
    if (x.b > x.e || y.b > y.e || z.b > z.e) return;
    const int ncx = (dx + C - 1) / C, ncy = (dy + C - 1) / C, ncz = (dz + C - 1) / C;
    const int cx0 = std::max(0, x.b / C), cy0 = std::max(0, y.b / C), cz0 = std::max(0, z.b / C);
    const int cx1 = std::min(ncx - 1, x.e / C), cy1 = std::min(ncy - 1, y.e / C), cz1 = std::min(ncz - 1, z.e / C);
    for (int cz = cz0; cz <= cz1; ++cz)
        for (int cy = cy0; cy <= cy1; ++cy)
            for (int cx = cx0; cx <= cx1; ++cx) out.push_back(((int64_t)cz * ncy + cy) * ncx + cx);
}

AxisRange ref_range_for_used_axis(AxisRange t, int dim)  // F5:1726-1749
{
    if (t.b > t.e) return t;
    int lo = t.b, hi = t.e;
    const int ib = std::max(t.b, 1), ie = std::min(t.e, dim - 3);
    if (ib <= ie) { lo = std::min(lo, ib - 1); hi = std::max(hi, ie + 2); }
    const int nb = std::max(t.b, 0), ne = std::min(t.e, dim - 2);
    if (nb <= ne) hi = std::max(hi, ne + 1);
    return {std::max(0, lo), std::min(dim - 1, hi)};
}
}  // namespace

void make_plan(Plan& pl, int nx, int ny, int nz, int px, int py, int pz, int chunk_high, int chunk_low, int ax,
               int ay, int az, int bx, int by, int bz)
{
    pl = Plan();
    const int mx = px / 2, my = py / 2, mz = pz / 2;
    const int lx = px / 4, ly = py / 4, lz = pz / 4;
    bool sl = false;
    const AxisRange qx = normalize_axis(ax, bx, nx, sl);  // F5:1637-1644
    const AxisRange qy = normalize_axis(ay, by, ny, sl);
    const AxisRange qz = normalize_axis(az, bz, nz, sl);
    pl.q = {qx.b, qx.e, qy.b, qy.e, qz.b, qz.e};
    pl.is_slice = sl;
    for (int b = 0; b < 7; ++b) {  // F5:1689-1711 build_residual_plan
        const int k = b + 1;
        const AxisRange rx = range_for_parity(qx, k & 1), ry = range_for_parity(qy, (k >> 1) & 1),
                        rz = range_for_parity(qz, (k >> 2) & 1);
        if (rx.b > rx.e || ry.b > ry.e || rz.b > rz.e) continue;
        BlockPlan& hp = pl.high[b];
        hp.valid = true;
        hp.r = {rx.b, rx.e, ry.b, ry.e, rz.b, rz.e};
        chunks_for_box(mx, my, mz, chunk_high, rx, ry, rz, hp.chunks);
        // F5:1751-1770 build_high_ref_region
        const AxisRange ux = (k & 1) ? ref_range_for_used_axis(rx, mx) : rx;
        const AxisRange uy = (k & 2) ? ref_range_for_used_axis(ry, my) : ry;
        const AxisRange uz = (k & 4) ? ref_range_for_used_axis(rz, mz) : rz;
        pl.regions.push_back({ux.b, ux.e, uy.b, uy.e, uz.b, uz.e});
    }
    {  // F5:1772-1782 build_query_sub0_region, appended as F5:2464-2469 does
        const AxisRange sx = range_for_parity(qx, 0), sy = range_for_parity(qy, 0), sz = range_for_parity(qz, 0);
        if (sx.b <= sx.e && sy.b <= sy.e && sz.b <= sz.e) pl.regions.push_back({sx.b, sx.e, sy.b, sy.e, sz.b, sz.e});
    }
    for (int b = 0; b < 7; ++b) {  // F5:1784-1830 build_low_residual_plan
        const int k = b + 1;
        std::set<int64_t> set;
        AxisRange ux{1, 0}, uy{1, 0}, uz{1, 0};
        for (const Box& R : pl.regions) {
            const AxisRange rx{R.x0, R.x1}, ry{R.y0, R.y1}, rz{R.z0, R.z1};
            if (rx.b > rx.e || ry.b > ry.e || rz.b > rz.e) continue;
            const AxisRange px_ = range_for_parity(rx, k & 1), py_ = range_for_parity(ry, (k >> 1) & 1),
                            pz_ = range_for_parity(rz, (k >> 2) & 1);
            if (px_.b > px_.e || py_.b > py_.e || pz_.b > pz_.e) continue;
            std::vector<int64_t> ch;
            chunks_for_box(lx, ly, lz, chunk_low, px_, py_, pz_, ch);
            set.insert(ch.begin(), ch.end());
            auto merge = [](AxisRange& d, AxisRange s) {  // F5:1713-1724
                if (s.b > s.e) return;
                if (d.b > d.e) { d = s; return; }
                d.b = std::min(d.b, s.b);
                d.e = std::max(d.e, s.e);
            };
            merge(ux, px_);
            merge(uy, py_);
            merge(uz, pz_);
        }
        if (set.empty()) continue;
        BlockPlan& lp = pl.low[b];
        lp.valid = true;
        lp.r = {ux.b, ux.e, uy.b, uy.e, uz.b, uz.e};
        lp.chunks.assign(set.begin(), set.end());
    }
}

}  // namespace stz
