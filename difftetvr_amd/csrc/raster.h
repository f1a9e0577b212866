// Fixed-point screen-tile rasterisation shared by the binning passes of the tile renderers: tile ranges of a
// screen triangle, the per-tile reject test and the exact 8x8 pixel-centre coverage (the tie rule gives each
// shared edge to exactly one side).
#pragma once
#include "dtv_device.h"
#include "dtv_prims.h"

namespace dtv {

struct FaceBins {
    int32_t tx0, tx1, ty0, ty1;  // tile range of the face's pixel-centre bounding box (empty: tx0 > tx1)
    int32_t shift;               // log2 of the tile edge in pixels (3, 4 or 5)
    ScreenTri S;
};

// Tile range of screen triangle S (its pixel-centre bounding box, clipped to the view).
__device__ __forceinline__ FaceBins screen_bins(const ViewUniforms& U, const ScreenTri& S) {
    FaceBins b;
    b.tx0 = 0;
    b.tx1 = -1;
    b.ty0 = 0;
    b.ty1 = -1;
    b.shift = U.tile_shift;
    b.S = S;
    int32_t x0, x1, y0, y1;
    tri_pixel_bbox(U, b.S, x0, x1, y0, y1);
    if (x0 > x1 || y0 > y1) return b;
    b.tx0 = x0 >> b.shift;
    b.tx1 = x1 >> b.shift;
    b.ty0 = y0 >> b.shift;
    b.ty1 = y1 >> b.shift;
    return b;
}

__device__ __forceinline__ FaceBins face_bins(const ViewUniforms& U, const ProjectedTet& T, int f) {
    ScreenTri S;
    face_screen(U, T, f, S);
    return screen_bins(U, S);
}

// Edge e of the oriented screen triangle: E(px, py) = dx (py - Y_a) - dy (px - X_a) in 1/256-pixel units; a
// pixel centre is inside when every E > 0, or E == 0 on a tie edge (dy < 0, or dy == 0 and dx > 0).
struct TileEdge {
    long long C;       // E at the centre of the tile's first pixel
    long long sx, sy;  // change of E per pixel step in x and y
    bool tie;
};

// (tx, ty): a tile of 1 << shift pixels (8 by default).
__device__ __forceinline__ TileEdge tile_edge(const ScreenTri& S, int e, int tx, int ty, int shift = 3) {
    const int a = e, b = e == 2 ? 0 : e + 1;
    const long long dx = S.X[b] - S.X[a], dy = S.Y[b] - S.Y[a];
    const long long px0 = 256LL * ((long long)tx << shift) + 128, py0 = 256LL * ((long long)ty << shift) + 128;
    TileEdge t;
    t.C = dx * (py0 - S.Y[a]) - dy * (px0 - S.X[a]);
    t.sx = -dy * 256;
    t.sy = dx * 256;
    t.tie = dy < 0 || (dy == 0 && dx > 0);
    return t;
}

// False when some edge function is negative at every pixel centre of tile (tx, ty): a record of that tile
// could not cover a pixel.  (The maximum of E over the tile's centre lattice sits at a corner.)
__device__ __forceinline__ bool tile_may_cover(const ScreenTri& S, int tx, int ty, int shift = 3) {
    const long long n1 = (1 << shift) - 1;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const TileEdge t = tile_edge(S, e, tx, ty, shift);
        const long long mx = t.C + (t.sx > 0 ? n1 * t.sx : 0) + (t.sy > 0 ? n1 * t.sy : 0);
        if (mx < 0) return false;
    }
    return true;
}

// Pixel (x, y)'s centre inside oriented screen triangle (X, Y): every edge function E > 0, or E == 0 on a tie
// edge; the rule of tile_coverage, for one pixel.
__device__ __forceinline__ bool pixel_inside(const int32_t X[3], const int32_t Y[3], int x, int y) {
    const long long px = 256LL * x + 128, py = 256LL * y + 128;
    bool in = true;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const int a = e, b = e == 2 ? 0 : e + 1;
        const long long dx = (long long)X[b] - X[a], dy = (long long)Y[b] - Y[a];
        const long long E = dx * (py - Y[a]) - dy * (px - X[a]);
        const bool tie = dy < 0 || (dy == 0 && dx > 0);
        in = in && (E > 0 || (E == 0 && tie));
    }
    return in;
}

// Faces whose bounding box spans more tiles than this keep every tile of it (bounded per-thread work).
constexpr int kMaxTestedTiles = 64;

__device__ __forceinline__ bool test_tiles(const FaceBins& b) {
    return (b.tx1 - b.tx0 + 1) * (b.ty1 - b.ty0 + 1) <= kMaxTestedTiles;
}

__device__ __forceinline__ uint32_t face_record_count(const FaceBins& b) {
    if (b.tx0 > b.tx1 || b.ty0 > b.ty1) return 0u;
    if (!test_tiles(b)) return (uint32_t)((b.tx1 - b.tx0 + 1) * (b.ty1 - b.ty0 + 1));
    uint32_t c = 0;
    for (int ty = b.ty0; ty <= b.ty1; ty++)
        for (int tx = b.tx0; tx <= b.tx1; tx++) c += tile_may_cover(b.S, tx, ty, b.shift) ? 1u : 0u;
    return c;
}

// Exact coverage of tile (tx, ty): bit ly * 8 + lx is set when the centre of pixel (8 tx + lx, 8 ty + ly) is
// inside the triangle; pixels outside the viewport are clear.  Edges whose extremes over the tile decide every
// pixel at once are not evaluated per pixel.
__device__ __forceinline__ uint64_t tile_coverage(const ScreenTri& S, int tx, int ty, int width, int height) {
    const int nx = min(8, width - 8 * tx), ny = min(8, height - 8 * ty);
    const uint64_t row = nx >= 8 ? 0xFFull : ((1ull << nx) - 1ull);
    uint64_t m = 0;
    for (int ly = 0; ly < ny; ly++) m |= row << (8 * ly);
#pragma unroll 1
    for (int e = 0; e < 3; e++) {
        const TileEdge t = tile_edge(S, e, tx, ty);
        const long long ex = 7 * t.sx, ey = 7 * t.sy;
        const long long emax = t.C + (ex > 0 ? ex : 0) + (ey > 0 ? ey : 0);
        const long long emin = t.C + (ex < 0 ? ex : 0) + (ey < 0 ? ey : 0);
        if (emax < 0 || (emax == 0 && !t.tie)) return 0ull;
        if (emin > 0 || (emin == 0 && t.tie)) continue;  // the edge admits every pixel of the tile
        uint64_t em = 0;
        // sx = -256 dy and sy = 256 dx, so E = 256 W + r with W = q + ly dx - lx dy, q = floor(C / 256),
        // r = C mod 256: E > 0 or (E == 0 on a tie edge)  <=>  W >= wmin.  The edge crosses the tile, so
        // emin <= 0 <= emax bounds |E| by 7 (|sx| + |sy|) and |W| by 7 (|dx| + |dy|) + 1 < 2^31 for screen
        // coordinates below 2^26 (the guard band keeps them below 2^25.2): exact in 32-bit integers.
        const long long dx = t.sy >> 8, dy = -(t.sx >> 8);
        if (dx > -(1ll << 27) && dx < (1ll << 27) && dy > -(1ll << 27) && dy < (1ll << 27)) {
            const int dxi = (int)dx, dyi = (int)dy;
            const int r = (int)(t.C & 255);
            const int wmin = (r > 0 || t.tie) ? 0 : 1;
            const int q = (int)(t.C >> 8);
            // only values at the tile's 64 pixel centres are formed (no step past the tile can overflow)
#pragma unroll
            for (int ly = 0; ly < 8; ly++) {
                const int Wr = q + ly * dxi;
                uint32_t rm = 0;
#pragma unroll
                for (int lx = 0; lx < 8; lx++) rm |= (uint32_t)(Wr - lx * dyi >= wmin) << lx;
                em |= (uint64_t)rm << (8 * ly);
            }
        } else {
            long long Er = t.C;
#pragma unroll 1
            for (int ly = 0; ly < 8; ly++) {
                long long E = Er;
                for (int lx = 0; lx < 8; lx++) {
                    if (E > 0 || (E == 0 && t.tie)) em |= 1ull << (8 * ly + lx);
                    E += t.sx;
                }
                Er += t.sy;
            }
        }
        m &= em;
        if (m == 0) return 0ull;
    }
    return m;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_and_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v &= __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v |= __shfl_xor(v, d, 64);
    return v;
}

}  // namespace dtv
