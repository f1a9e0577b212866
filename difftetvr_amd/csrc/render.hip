// Tile renderer and its adjoint on gfx950.
//
// One wave64 per 8x8 screen tile.  The tile's sorted (tile, face) records are consumed in batches of 32: each
// lane stages one record in LDS (tet vertices, colours, face planes, fixed-point edge functions of the record's
// face); every lane (pixel) then tests all 32 faces against its pixel centre, giving a 32-bit coverage mask.
// A wave prefix sum turns the masks into a dense list of (pixel, record) fragments in LDS, and the expensive
// per-fragment work (ray-tet intersection, barycentric colours, four sub-segment integrals; in the adjoint also
// their derivatives) runs over that list with every lane busy, in rounds of 128.  Only the cheap per-pixel
// compositing (or, in the adjoint, the inversion of the 'over' operator) walks each pixel's fragments in order.
//
// Reference maths (Data/Shaders/...):
//   ray through pixel centre           Intersect/IntersectRasterization.glsl:141-144
//   ray-tet entry/exit                 Intersect/RayIntersectionTests.glsl:78-106
//   barycentric colours                Common/BarycentricInterpolation.glsl:29-38
//   4 sub-segments, accumulateLinearConst  IntersectRasterization.glsl:228-241, Common/ForwardCommon.glsl:42-49
//   early ray termination              PPLL/Forward.glsl:168-173
// The adjoint re-traverses each pixel back to front from the stored final colour, inverting the 'over'
// operator per tet (the inversion of AdjointIntersectRasterization.glsl:213-228) and resetting to stored
// checkpoints (every K fragments, and before any fragment that takes the transmittance below 1/8 of the last
// checkpoint's) to bound inversion drift.  Gradients of the 4 tet vertices are reduced per record in LDS before one global
// atomic per value.
#include <map>
#include <mutex>

#include "render_common.h"

namespace dtv {

constexpr int kRoundDefault = 128;          // fragments shaded per round of the forward (segments per pass)
constexpr int kRoundAdj = 64;               // fragments per round in the adjoint: one per lane
constexpr int kMaxFrags = kBatch * 64;      // fragments a batch can produce
constexpr float kMissMarker = -1.0f;        // transmittance slot of a fragment whose ray missed the tet
constexpr float kTinyT = 0x1p-100f;         // transmittance below which a fragment's input is not recovered

// The forward's fragment list and round values (kRound fragments shaded per round).
template <int kRound>
struct FragListLds {
    uint16_t entry[kMaxFrags];  // (pixel lane << 5) | record index
    float4 val[kRound];         // fragment colour F
    float rd[3][64];            // ray direction of each pixel
};

// The adjoint's: with the records, 13 KB per wave, so that three waves fit in each SIMD (a CU has 160 KB).  A
// batch whose fragments exceed kAdjFrags is processed in two parts of 16 records; the per-fragment gradient rows
// are reduced half a round (32 fragments) at a time; ray directions stay in registers and move between lanes by
// permute.  The flush's vertex hash table spans entry and w, which are free by then.
constexpr int kAdjFrags = 1024;
struct AdjFragLds {
    uint16_t entry[kAdjFrags];  // (pixel lane << 5) | record index, one part of a batch
    float w[32][29];            // gradient rows of half a round, one per lane (29: conflict-free column access)
    float4 val[kRoundAdj];      // F, then dL/dF (.w: d/d(alpha), alpha = 1 - F.w); w = kMissMarker on a miss
};

// Ray direction of pixel lane pl from the lanes' own (register) directions; every lane must take part.
__device__ __forceinline__ float3_ lane_ray(float3_ mine, int pl) {
    return f3(__shfl(mine.x, pl, 64), __shfl(mine.y, pl, 64), __shfl(mine.z, pl, 64));
}

// kCell: renders the items with per-cell colours (BatchItem::cell_col), else those with per-vertex colours; a
// batch that holds both is rendered by one launch of each, every tile by the one matching its item.
// kRound: fragments shaded per round (64, 128 or 256: segments per pass).  kSub: binning tiles larger than 8x8
// (the workgroup's 8x8 sub-tile tests its own coverage).
template <bool kCell, int kRound, bool kSub>
__device__ __forceinline__ void forward_tile(const BatchItem* __restrict__ items, int nitems,
                                             const uint4* __restrict__ recs, const uint2* __restrict__ ranges,
                                             const PixelState& ps_all, uint32_t gsub, RecordLds& L,
                                             FragListLds<kRound>& F) {
    const BatchItem& B = items[item_of_sub(items, nitems, gsub)];
    if ((B.cell_col != nullptr) != kCell) return;
    const ViewUniforms& U = B.U;
    const float* __restrict__ pos = B.pos;
    const float* __restrict__ col = kCell ? B.cell_col : B.col;
    float4* __restrict__ out_image = reinterpret_cast<float4*>(B.image);
    const PixelState ps = item_pixel_state(ps_all, B);
    const SubTile st = sub_tile(B, gsub);
    const uint32_t gtile = st.gtile;
    const int lane = threadIdx.x;
    const uint64_t lt = lanemask_lt();
    const int lx = lane & 7, ly = lane >> 3;
    const int x = st.x0 + lx, y = st.y0 + ly;
    // binning tiles larger than 8x8: coverage is tested here, from face triangles stashed in the fragment
    // list's space (free while a batch is staged)
    int32_t* stash = kSub ? reinterpret_cast<int32_t*>(F.entry) : nullptr;
    const bool valid = x < U.width && y < U.height;
    const float3_ ro = f3(U.camPos[0], U.camPos[1], U.camPos[2]);
    {
        const float3_ rd = ray_direction(U, x, y);
        F.rd[0][lane] = rd.x;
        F.rd[1][lane] = rd.y;
        F.rd[2][lane] = rd.z;
    }
    const float att = U.attenuation;
    const uint2 rg = ranges[gtile];
    const int P = U.width * U.height;
    const int pix = valid ? pixel_index(U, x, y) : 0;

    float4 C = make_float4(0.f, 0.f, 0.f, 1.0f);  // premultiplied RGB, transmittance
    uint32_t n = 0, last = rg.x, nck = 0, nlast = 0;
    bool done = !valid;
    // Two-stage prefetch: while a batch is shaded, the next batch's vertex data and the record words of the
    // batch after it are in flight.
    RecordIn cur, nxt;
    RecordData cd;
    if (lane < kBatch && rg.x + lane < rg.y) {
        cur = load_record(recs, rg.x + lane);
        fetch_record_data<kCell>(pos, col, cur, cd);
    }
    if (lane < kBatch && rg.x + kBatch + lane < rg.y) nxt = load_record(recs, rg.x + kBatch + lane);
    for (uint32_t base = rg.x; base < rg.y; base += kBatch) {
        const int cnt = (int)min((uint32_t)kBatch, rg.y - base);
        if (lane < cnt) store_record<kSub>(U, cur, cd, L, lane, stash);
        if (lane < kBatch && base + kBatch + lane < rg.y) {
            cur = nxt;
            fetch_record_data<kCell>(pos, col, cur, cd);
        }
        if (lane < kBatch && base + 2 * kBatch + lane < rg.y) nxt = load_record(recs, base + 2 * kBatch + lane);
        __syncthreads();
        if constexpr (kSub) {
            subtile_masks(L, stash, cnt, x, y, valid, lane);
            __syncthreads();
        }
        const uint32_t mask = done ? 0u : lane_coverage(L, cnt, lane);
        // the records' coverage masks are consumed: their slots hold the list's rank masks
        uint64_t* rank_mask = reinterpret_cast<uint64_t*>(&L.mask[0][0]);
        const uint32_t total = build_fragment_list(F.entry, rank_mask, mask, false, lane);
        const uint32_t nmine = (uint32_t)__popc(mask);
        RankCursor cur_e;
        __syncthreads();
        for (uint32_t r0 = 0; r0 < total; r0 += kRound) {
            for (int i = lane; i < kRound; i += 64) {
                const uint32_t e = r0 + i;
                if (e >= total) break;
                const uint32_t ent = F.entry[e];
                const int pl = ent >> 5, j = ent & 31;
                const float3_ rdp = f3(F.rd[0][pl], F.rd[1][pl], F.rd[2][pl]);
                FragEval E;
                eval_fragment<kCell, false>(L, j, ro, rdp, att, seg_scale(U, rdp), E);
                F.val[i] = E.hit ? E.F : make_float4(0.f, 0.f, 0.f, kMissMarker);
            }
            __syncthreads();
            uint32_t e;
            while (!done && cur_e.next(rank_mask, nmine, lt, r0 + kRound, e)) {
                if (U.use_ert && (1.0f - C.w) > U.ert_alpha) {
                    done = true;
                    break;
                }
                const float4 f = F.val[e - r0];
                // C.w is the transmittance T, the product of the fragments' transmittances (relative precision at
                // any depth), so the adjoint recovers each fragment's input as T_out / T_f.  Only an underflow
                // (T_out below kTinyT) cannot be inverted: a checkpoint goes before the fragment that causes it.
                // The state before the first fragment (T = 1) needs no slot.  A missed fragment's value is
                // (0, 0, 0, kMissMarker): with transmittance 1 it leaves the state as it is, without a branch.
                const float Tout = C.w * (f.w != kMissMarker ? f.w : 1.0f);
                if (C.w >= kTinyT && Tout < kTinyT && nlast != n && (int)nck < U.max_ckpt) {
                    ps.ckpt_t[(int64_t)nck * P + pix] = C.w;
                    ps.ckpt_n[(int64_t)nck * P + pix] = n;
                    nck++;
                    nlast = n;
                }
                {
#pragma clang fp contract(fast)
                    C.x = C.x + C.w * f.x;
                    C.y = C.y + C.w * f.y;
                    C.z = C.z + C.w * f.z;
                }
                C.w = Tout;
                n++;
                last = base + (F.entry[e] & 31) + 1;
                // periodic checkpoints bound the rounding the adjoint's chain of divisions accumulates; the last
                // slot is kept for the underflow rule above
                if ((int)(n - nlast) >= U.ckpt_interval && C.w >= kTinyT && (int)nck < U.max_ckpt - 1) {
                    ps.ckpt_t[(int64_t)nck * P + pix] = C.w;
                    ps.ckpt_n[(int64_t)nck * P + pix] = n;
                    nck++;
                    nlast = n;
                }
            }
            __syncthreads();
        }
        if (ballot(!done) == 0ull) break;
    }
    if (valid) {
        const float T = C.w;
        out_image[pix] = make_float4(C.x + T * U.clear[0], C.y + T * U.clear[1], C.z + T * U.clear[2],
                                     (1.0f - T) + T * U.clear[3]);
        ps.ray_color[pix] = C;
        ps.frag_count[pix] = n;
        ps.last_rec[pix] = last;
        ps.num_ckpt[pix] = nck;
    }
    // the tile's composited fragments: the adjoint's launch order ranks tiles by them
    const uint32_t tile_total = wave_inclusive_scan(n);
    if (lane == 63) ps.tile_frags[gsub] = tile_total;
}

// Next entry of the launch order for a persistent workgroup: one claim per tile from the launch's queue
// counter (zeroed before the launch), uniform across the wave.  Every tile is claimed, none assigned by block
// index: a workgroup the GPU cannot hold at first (the occupancy estimate is an upper bound) starts only when
// another exits, and must not be holding a tile meanwhile.
__device__ __forceinline__ uint32_t claim_tile(uint32_t* __restrict__ queue) {
    uint32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(queue, 1u);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl((int)t, 0, 64));
}

// kQueue = false: one workgroup per tile of the launch order.  kQueue = true: as many workgroups as the GPU
// holds at once, each claiming the next tile of the launch order from `queue` until all ntiles are taken.
template <bool kCell, int kRound, bool kSub, bool kQueue>
__global__ __launch_bounds__(64) void k_render_forward(const BatchItem* __restrict__ items, int nitems,
                                                        const uint32_t* __restrict__ order,
                                                        const uint4* __restrict__ recs,
                                                        const uint2* __restrict__ ranges, PixelState ps_all,
                                                        uint32_t* __restrict__ queue, uint32_t ntiles) {
    __shared__ RecordLds L;
    __shared__ FragListLds<kRound> F;
    if constexpr (!kQueue) {
        forward_tile<kCell, kRound, kSub>(items, nitems, recs, ranges, ps_all, global_tile_of_block(order), L, F);
    } else {
        for (uint32_t t = claim_tile(queue); t < ntiles; t = claim_tile(queue)) {
            forward_tile<kCell, kRound, kSub>(items, nitems, recs, ranges, ps_all,
                                              order != nullptr ? order[t] : t, L, F);
            __syncthreads();
        }
    }
}

// kRef = false: exact derivative of the forward.  kRef = true: the reference's adjoint as written
// (AdjointIntersectRasterization.glsl:111-362): per pixel from the blended output image, back to front, the
// 'over' operator is inverted per sub-segment 3, 2, 1 (sub-segment 0 is skipped) with no checkpoints, and
// each fragment's parameter gradients follow frag_backward_reference.
// kCell: the items with per-cell colours (exact adjoint only), kSub: binning tiles larger than 8x8, as
// k_render_forward.
template <bool kRef, bool kCell, bool kSub>
__device__ __forceinline__ void adjoint_tile(const BatchItem* __restrict__ items,
                                             const AdjointTarget* __restrict__ targets, int nitems,
                                             const uint4* __restrict__ recs, const uint2* __restrict__ ranges,
                                             const PixelState& ps_all, int use_abs, uint32_t gsub, RecordLds& L,
                                             AdjFragLds& A, float4 (&s_seg)[kRef ? 3 : 1][kRoundAdj]) {
    const int item = item_of_sub(items, nitems, gsub);
    const BatchItem& B = items[item];
    if ((B.cell_col != nullptr) != kCell) return;
    const ViewUniforms& U = B.U;
    const float* __restrict__ pos = B.pos;
    const float* __restrict__ col = kCell ? B.cell_col : B.col;
    const float4* __restrict__ dimage = reinterpret_cast<const float4*>(targets[item].dimage);
    if (dimage == nullptr) return;  // item not differentiated
    // positions and vertex colours accumulate in the item's interleaved rows (8 floats per vertex: xyz, rgba)
    float* __restrict__ vacc = targets[item].acc;
    float* __restrict__ grad_col = targets[item].grad_col;  // per-cell colours: straight to the tet
    const PixelState ps = item_pixel_state(ps_all, B);
    const SubTile st = sub_tile(B, gsub);
    const uint32_t gtile = st.gtile;
    const int lane = threadIdx.x;
    const int lx = lane & 7, ly = lane >> 3;
    const int x = st.x0 + lx, y = st.y0 + ly;
    const bool valid = x < U.width && y < U.height;
    int32_t* stash = kSub ? reinterpret_cast<int32_t*>(A.entry) : nullptr;
    const float3_ ro = f3(U.camPos[0], U.camPos[1], U.camPos[2]);
    const float3_ my_rd = ray_direction(U, x, y);
    const float att = U.attenuation;
    const uint2 rg = ranges[gtile];
    const int P = U.width * U.height;
    const int pix = valid ? pixel_index(U, x, y) : 0;

    uint32_t natom = 0;  // gradient atomics this lane issued (counted when ps.counters is set)
    // exact mode: C.w = the transmittance behind the fragment being undone (only it is needed); reference mode:
    // the blended premultiplied RGBA the reference inverts
    float4 C = make_float4(0.f, 0.f, 0.f, 1.0f);
    float dC[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t n = 0, last = 0;
    int jck = -1;                // next checkpoint slot (back to front)
    // Checkpoint jck's fragment count and transmittance, and the count of the one before it, are held in
    // registers: the next checkpoint is always loaded well before it is needed.
    uint32_t ck_n = 0xFFFFFFFFu, ck_n2 = 0xFFFFFFFFu;
    float ck_t = 1.0f;
    if (valid) {
        C = ps.ray_color[pix];
        n = ps.frag_count[pix];
        last = ps.last_rec[pix];
        jck = (int)ps.num_ckpt[pix] - 1;
        if (jck >= 0) {
            ck_n = ps.ckpt_n[(int64_t)jck * P + pix];
            ck_t = ps.ckpt_t[(int64_t)jck * P + pix];
            if (jck >= 1) ck_n2 = ps.ckpt_n[(int64_t)(jck - 1) * P + pix];
        }
        const float4 di = dimage[pix];
        dC[0] = di.x;
        dC[1] = di.y;
        dC[2] = di.z;
        if constexpr (kRef) {
            // the reference starts from the blended output image and the raw image adjoint
            const float T = C.w;
            C = make_float4(C.x + T * U.clear[0], C.y + T * U.clear[1], C.z + T * U.clear[2],
                            (1.0f - T) + T * U.clear[3]);
            dC[3] = di.w;
            jck = -1;
        } else {
            dC[3] = di.w * (1.0f - U.clear[3]) - (di.x * U.clear[0] + di.y * U.clear[1] + di.z * U.clear[2]);
        }
    }
    bool done = !valid || n == 0;
    if (ballot(!done) == 0ull) return;
    const uint32_t nrec = rg.y - rg.x;
    const int nbatch = (int)((nrec + kBatch - 1) / kBatch);
    // Batches run back to front.  The next batch's record words are loaded when the current batch starts and its
    // vertex data just before the current batch's gradient atomics, so that staging it never waits on memory.
    RecordIn nxt;
    RecordData nd;
    {
        const uint32_t b0 = rg.x + (uint32_t)(nbatch - 1) * kBatch;
        if (lane < kBatch && b0 + lane < rg.y) {
            nxt = load_record(recs, b0 + lane);
            fetch_record_data<kCell>(pos, col, nxt, nd);
        }
    }
    for (int bi = nbatch - 1; bi >= 0; bi--) {
        const uint32_t base = rg.x + (uint32_t)bi * kBatch;
        const int cnt = (int)min((uint32_t)kBatch, rg.y - base);
        if (lane < cnt) store_record<kSub>(U, nxt, nd, L, lane, stash);
        if (bi > 0 && lane < kBatch) nxt = load_record(recs, base - kBatch + lane);
        // lanes j and j + 32 own record j's 28 values (first / second half of 14) for this batch
        float acc[14];
#pragma unroll
        for (int q = 0; q < 14; q++) acc[q] = 0.f;
        __syncthreads();
        if constexpr (kSub) {
            subtile_masks(L, stash, cnt, x, y, valid, lane);
            __syncthreads();
        }
        // records at or after `last` were not composited in the forward (early ray termination)
        const uint32_t nlive = last > base ? min((uint32_t)cnt, last - base) : 0u;
        const uint32_t live = nlive >= 32u ? 0xFFFFFFFFu : ((1u << nlive) - 1u);
        const uint32_t mask_all = done ? 0u : (lane_coverage(L, cnt, lane) & live);
        // parts (back to front): all 32 records, or records 16-31 then 0-15 when the list exceeds kAdjFrags
        int nparts = 1;
        {
            const uint32_t inc = wave_inclusive_scan((uint32_t)__popc(mask_all));
            if ((uint32_t)__builtin_amdgcn_readlane((int)inc, 63) > (uint32_t)kAdjFrags) nparts = 2;
        }
        for (int part = 0; part < nparts; part++) {
        const uint32_t pmask = nparts == 1 ? 0xFFFFFFFFu : (part == 0 ? 0xFFFF0000u : 0x0000FFFFu);
        const uint32_t mask = done ? 0u : (mask_all & pmask);
        uint32_t total;
        const uint32_t off = build_fragment_list_by_pixel(A.entry, mask, true, lane, total);
        const uint32_t mine_end = off + (uint32_t)__popc(mask);
        __syncthreads();
        for (uint32_t r0 = 0; r0 < total; r0 += kRoundAdj) {
            // A: fragment colours (one fragment per lane; E stays in registers for phase C)
            FragEval E;
            E.hit = false;
            {
                const uint32_t e = r0 + lane;
                const uint32_t ent = e < total ? A.entry[e] : (uint32_t)(lane << 5);
                const int pl = ent >> 5, j = ent & 31;
                const float3_ rdp = lane_ray(my_rd, pl);
                if (e < total) {
                    eval_fragment<kCell, true, kRef>(L, j, ro, rdp, att, seg_scale(U, rdp), E);
                    A.val[lane] = E.hit ? E.F : make_float4(0.f, 0.f, 0.f, kMissMarker);
                    if constexpr (kRef) {
#pragma unroll
                        for (int s = 1; s < 4; s++) s_seg[s - 1][lane] = E.seg[s];
                    }
                }
            }
            __syncthreads();
            // B: per pixel, back to front: undo the composite, dL/dF of each fragment
            const uint32_t e0 = max(off, r0), e1 = min(mine_end, r0 + kRoundAdj);
            for (uint32_t e = e0; e < e1; e++) {
                if constexpr (kRef) {
                    // reference: invert sub-segments 3..1 of the fragment; dL/d(sub-segment colour) replaces
                    // the sub-segment colour in s_seg (AdjointIntersectRasterization.glsl:204-228)
                    if (A.val[e - r0].w != kMissMarker) {
#pragma unroll
                        for (int s = 3; s > 0; s--) {
                            const float4 acc = s_seg[s - 1][e - r0];
                            const float ain = (acc.w - C.w) / (acc.w - 1.0f);
                            const float4 cin = make_float4(C.x - (1.0f - ain) * acc.x, C.y - (1.0f - ain) * acc.y,
                                                           C.z - (1.0f - ain) * acc.z, ain);
                            s_seg[s - 1][e - r0] = make_float4((1.0f - ain) * dC[0], (1.0f - ain) * dC[1],
                                                               (1.0f - ain) * dC[2], (1.0f - ain) * dC[3]);
                            dC[3] = dC[3] * (1.0f - acc.w) - (dC[0] * acc.x + dC[1] * acc.y + dC[2] * acc.z);
                            C = cin;
                        }
                    }
                    n--;
                    continue;
                }
                if (jck >= 0 && ck_n == n) {
                    C.w = ck_t;
                    jck--;
                    ck_n = ck_n2;
                    if (jck >= 0) {
                        ck_t = ps.ckpt_t[(int64_t)jck * P + pix];
                        ck_n2 = jck >= 1 ? ps.ckpt_n[(int64_t)(jck - 1) * P + pix] : 0xFFFFFFFFu;
                    }
                }
                const float4 f = A.val[e - r0];
                // the fragment's input transmittance: a checkpoint, 1 before the first fragment, or the output
                // divided by the fragment's transmittance (a corrected quotient: the rounding compounds along the
                // chain back to the next checkpoint).  Below kTinyT the forward's rule leaves no checkpoint only
                // where the input was below kTinyT too: that fragment's contribution (at most kTinyT |dL/dI|) is
                // dropped.  Formed without branches; a missed fragment (f.w = kMissMarker) keeps the state, and
                // its slot's value is not read.
                const bool hit = f.w != kMissMarker;
                float Tin = C.w < kTinyT ? 0.0f : div_nr(C.w, f.w);
                Tin = n == 1 ? 1.0f : Tin;
                Tin = (jck >= 0 && ck_n == n - 1) ? ck_t : Tin;
                A.val[e - r0] = make_float4(Tin * dC[0], Tin * dC[1], Tin * dC[2], Tin * dC[3]);
                {
#pragma clang fp contract(fast)
                    const float d3 = dC[3] * f.w - (dC[0] * f.x + dC[1] * f.y + dC[2] * f.z);
                    dC[3] = hit ? d3 : dC[3];
                }
                C.w = hit ? Tin : C.w;
                n--;
            }
            if (n == 0) done = true;
            __syncthreads();
            // C: gradients of the fragments (one per lane).  Each fragment's 28 values are stored into A.w (half
            // a round at a time); the owner lanes of each record then add its rows.
            int key = kBatch;  // record index of this lane's fragment, kBatch = none
            float g[28];
#pragma unroll
            for (int q = 0; q < 28; q++) g[q] = 0.f;
            {
                const uint32_t e = r0 + lane;
                const uint32_t ent = e < total ? A.entry[e] : (uint32_t)(lane << 5);
                const int pl = ent >> 5, j = ent & 31;
                const float3_ rdp = lane_ray(my_rd, pl);
                if (e < total && E.hit) {
                    if constexpr (kRef) {
                        const float4 dacc[3] = {s_seg[0][lane], s_seg[1][lane], s_seg[2][lane]};
                        frag_backward_reference(E, L, j, ro, rdp, att, dacc, use_abs != 0, g);
                    } else if constexpr (kCell) {
                        // positions in the usual slots; the tet's colour gradient in vertex 0's colour slots
                        float lam0[4], lam1[4], dcell[4];
                        float3_ Q0, Q1;
                        frag_backward_cell(E, L, j, rdp, att, seg_scale(U, rdp), A.val[lane], lam0, lam1, dcell, Q0,
                                           Q1);
                        if (use_abs) {
#pragma unroll
                            for (int ch = 0; ch < 4; ch++) dcell[ch] = fabsf(dcell[ch]);
                            Q0 = f3(fabsf(Q0.x), fabsf(Q0.y), fabsf(Q0.z));
                            Q1 = f3(fabsf(Q1.x), fabsf(Q1.y), fabsf(Q1.z));
                        }
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const float l0 = lam0[v], l1 = lam1[v];
                            g[7 * v + 0] = l0 * Q0.x + l1 * Q1.x;
                            g[7 * v + 1] = l0 * Q0.y + l1 * Q1.y;
                            g[7 * v + 2] = l0 * Q0.z + l1 * Q1.z;
                        }
#pragma unroll
                        for (int ch = 0; ch < 4; ch++) g[3 + ch] = dcell[ch];
                    } else {
                        float lam0[4], lam1[4], dcol0[4], dcol1[4];
                        float3_ Q0, Q1;
                        frag_backward_exact(E, L, j, rdp, att, seg_scale(U, rdp), A.val[lane], lam0, lam1, dcol0,
                                            dcol1, Q0, Q1);
                        if (use_abs) {
#pragma unroll
                            for (int ch = 0; ch < 4; ch++) {
                                dcol0[ch] = fabsf(dcol0[ch]);
                                dcol1[ch] = fabsf(dcol1[ch]);
                            }
                            Q0 = f3(fabsf(Q0.x), fabsf(Q0.y), fabsf(Q0.z));
                            Q1 = f3(fabsf(Q1.x), fabsf(Q1.y), fabsf(Q1.z));
                        }
#pragma unroll
                        for (int v = 0; v < 4; v++) {
                            const float l0 = lam0[v], l1 = lam1[v];
                            g[7 * v + 0] = l0 * Q0.x + l1 * Q1.x;
                            g[7 * v + 1] = l0 * Q0.y + l1 * Q1.y;
                            g[7 * v + 2] = l0 * Q0.z + l1 * Q1.z;
#pragma unroll
                            for (int ch = 0; ch < 4; ch++) g[7 * v + 3 + ch] = l0 * dcol0[ch] + l1 * dcol1[ch];
                        }
                    }
                    key = j;
                }
            }
            // Per half round (the fragments of lanes 0-31, then 32-63): each fragment stores its 28 values in
            // the row of its lane; the owner lanes of record j (j and j + 32) find the record's rows from a mask of
            // the lanes whose key equals j, formed from five ballots of the key bits, and add them.
#pragma unroll 1
            for (int h = 0; h < 2; h++) {
                const bool mine = (lane >> 5) == h && key < kBatch;
                const uint32_t jj = (uint32_t)(lane & 31);
                uint32_t eqj = (uint32_t)(ballot(mine) >> (32 * h));
#pragma unroll
                for (int b = 0; b < 5; b++) {
                    const uint32_t ones = (uint32_t)(ballot(mine && ((key >> b) & 1)) >> (32 * h));
                    eqj &= ((jj >> b) & 1u) ? ones : ~ones;
                }
                if (mine) {
#pragma unroll
                    for (int q = 0; q < 28; q++) A.w[lane & 31][q] = g[q];
                }
                __syncthreads();
                {
                    // two rows per step: 28 independent LDS reads in flight instead of 14
                    const int half = lane >> 5;
                    uint32_t m = eqj;
                    while (m) {
                        const int r1 = __ffs(m) - 1;
                        m &= m - 1u;
                        if (m) {
                            const int r2 = __ffs(m) - 1;
                            m &= m - 1u;
                            float a[14], b[14];
#pragma unroll
                            for (int q = 0; q < 14; q++) {
                                a[q] = A.w[r1][half * 14 + q];
                                b[q] = A.w[r2][half * 14 + q];
                            }
#pragma unroll
                            for (int q = 0; q < 14; q++) acc[q] += a[q] + b[q];
                        } else {
#pragma unroll
                            for (int q = 0; q < 14; q++) acc[q] += A.w[r1][half * 14 + q];
                        }
                    }
                }
                __syncthreads();
            }
        }
        }  // parts
        // Flush.  Lanes j and j + 32 hold record j's tet vertices 0, 1 and 2, 3 (7 values each): 128 vertex
        // slots, slot s = v * 32 + j.  Depth-adjacent tets share most vertices (typically 6-7 slots per distinct
        // vertex in a batch), so the slots of one vertex are summed before the global atomics: each slot finds
        // its vertex's entry in an LDS hash table keyed by vertex id (compare-and-swap; the first slot of a
        // vertex creates the entry), stores its values in its own LDS row and links the row into the entry's
        // list with one exchange; the entry's creator then sums the list.  (LDS float atomics, one per value and
        // slot, cost about 2.5 cycles per lane.)
        {
            const int j = lane & 31, half = lane >> 5;
            constexpr int kTab = 160;
            // over entry and w, free after the rounds: keys[kTab], head[kTab], rows[128][8] (7 values, next slot)
            int32_t* keys = reinterpret_cast<int32_t*>(&A.entry[0]);
            int32_t* head = keys + kTab;
            float* rows = reinterpret_cast<float*>(head + kTab);
            static_assert(sizeof(int32_t) * 2 * kTab + sizeof(float) * 8 * 128 <= sizeof(A.entry) + sizeof(A.w),
                          "flush tables exceed the fragment list space");
            for (int e = lane; e < kTab; e += 64) {
                keys[e] = -1;
                head[e] = -1;
            }
            __syncthreads();
            const int s0 = (2 * half) * 32 + j, s1 = s0 + 32;
            const int32_t* vids = &L.vid[0][0];
            const bool slot_live = j < cnt;
            if constexpr (kCell) {
                // the record's tet colour gradient (lane j holds it in vertex 0's colour slots) goes straight
                // to the tet; only positions are merged by vertex below
                if (half == 0 && slot_live) {
                    const int64_t t = L.meta[j] >> 4;
#pragma unroll
                    for (int c = 0; c < 4; c++) {
                        if (acc[3 + c] != 0.f) {
                            atomic_add_f32(&grad_col[4 * t + c], acc[3 + c]);
                            natom++;
                        }
                        acc[3 + c] = 0.f;
                    }
                }
            }
            const int32_t v0 = slot_live ? vids[s0] : 0, v1 = slot_live ? vids[s1] : 0;
            int ent[2] = {-1, -1};
            bool owner[2] = {false, false};
#pragma unroll
            for (int k = 0; k < 2; k++) {
                if (!slot_live) continue;
                const int32_t v = k == 0 ? v0 : v1;
                int e = (int)(((uint32_t)v * 2654435761u) >> 24) * kTab >> 8;  // multiplicative hash -> [0, kTab)
                for (int probe = 0; probe < kTab; probe++) {
                    const int32_t old = atomicCAS(&keys[e], -1, v);
                    if (old == -1 || old == v) {
                        ent[k] = e;
                        owner[k] = old == -1;
                        break;
                    }
                    e = e + 1 == kTab ? 0 : e + 1;
                }
            }
#pragma unroll
            for (int k = 0; k < 2; k++) {
                if (ent[k] < 0) continue;
                const int sl = k == 0 ? s0 : s1;
#pragma unroll
                for (int c = 0; c < 7; c++) rows[8 * sl + c] = acc[7 * k + c];
                rows[8 * sl + 7] = __int_as_float(atomicExch(&head[ent[k]], sl));
            }
            __syncthreads();
            // each entry's creator sums its list into its own row
#pragma unroll
            for (int k = 0; k < 2; k++) {
                if (!owner[k]) continue;
                float sum[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                for (int sl = head[ent[k]]; sl >= 0;) {
#pragma unroll
                    for (int c = 0; c < 7; c++) sum[c] += rows[8 * sl + c];
                    sl = __float_as_int(rows[8 * sl + 7]);
                }
                const int own = k == 0 ? s0 : s1;
#pragma unroll
                for (int c = 0; c < 7; c++) rows[8 * own + c] = sum[c];
            }
            // The creators' rows are listed compactly; then eight lanes per row add its seven sums to the
            // vertex's 32-byte row of the interleaved accumulator, one wave instruction per eight vertices.
            int32_t* list = reinterpret_cast<int32_t*>(&A.val[0]);  // free after the rounds: 256 words
            const uint64_t own0 = ballot(owner[0]), own1 = ballot(owner[1]);
            const uint32_t n0 = (uint32_t)popc64(own0), nown = n0 + (uint32_t)popc64(own1);
            if (owner[0]) list[popc64(own0 & lanemask_lt())] = s0;
            if (owner[1]) list[n0 + popc64(own1 & lanemask_lt())] = s1;
            __syncthreads();
            // The next batch's vertex data is requested before the atomics, so staging it does not wait for them.
            // Only lanes 0..31 hold that batch's record words.
            if (bi > 0 && lane < kBatch) fetch_record_data<kCell>(pos, col, nxt, nd);
            uint32_t issued = 0;
            {
                const int c = lane & 7;
                for (uint32_t b0 = 0; b0 < nown; b0 += 8) {
                    const uint32_t li = b0 + (uint32_t)(lane >> 3);
                    if (li < nown && c < 7) {
                        const int sl = list[li];
                        const float val = rows[8 * sl + c];
                        if (val != 0.f) {
                            atomic_add_f32(&vacc[8 * (int64_t)vids[sl] + c], val);
                            issued++;
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < 2; k++) {
                // a slot that found no entry (not possible with kTab > 128 ids) issues its own values
                if (!slot_live || ent[k] >= 0) continue;
                const int64_t g = k == 0 ? v0 : v1;
#pragma unroll
                for (int c = 0; c < 7; c++) {
                    const float val = acc[7 * k + c];
                    if (val == 0.f) continue;
                    issued++;
                    atomic_add_f32(&vacc[8 * g + c], val);
                }
            }
            natom += issued;
        }
        if (ballot(!done) == 0ull) break;
        __syncthreads();
    }
    if (ps.counters != nullptr) {
        const uint32_t tot = wave_inclusive_scan(natom);
        if (lane == 63 && tot) atomicAdd(ps.counters, (unsigned long long)tot);
    }
}

// Waves per SIMD: 3 for the exact adjoint (168 VGPRs of the 512 per lane, 13 KB of LDS), 2 for the reference
// one, whose sub-segment arrays need more of both.  kQueue as in k_render_forward.
template <bool kRef, bool kCell, bool kSub, bool kQueue>
__global__ __launch_bounds__(64, kRef ? 2 : 3) void k_render_adjoint(const BatchItem* __restrict__ items,
                                                        const AdjointTarget* __restrict__ targets, int nitems,
                                                        const uint32_t* __restrict__ order,
                                                        const uint4* __restrict__ recs,
                                                        const uint2* __restrict__ ranges, PixelState ps_all,
                                                        int use_abs, uint32_t* __restrict__ queue,
                                                        uint32_t ntiles) {
    __shared__ RecordLds L;
    __shared__ AdjFragLds A;
    __shared__ float4 s_seg[kRef ? 3 : 1][kRoundAdj];  // reference mode: sub-segment colours, then their adjoints
    if constexpr (!kQueue) {
        adjoint_tile<kRef, kCell, kSub>(items, targets, nitems, recs, ranges, ps_all, use_abs,
                                        global_tile_of_block(order), L, A, s_seg);
    } else {
        for (uint32_t t = claim_tile(queue); t < ntiles; t = claim_tile(queue)) {
            adjoint_tile<kRef, kCell, kSub>(items, targets, nitems, recs, ranges, ps_all, use_abs,
                                            order != nullptr ? order[t] : t, L, A, s_seg);
            __syncthreads();
        }
    }
}

// Ordered fragment export (tet*4 + face per covered record), up to max_per_pixel per pixel.
__global__ __launch_bounds__(64) void k_export_fragments(const BatchItem* __restrict__ items, int nitems,
                                                          const uint4* __restrict__ recs,
                                                          const uint2* __restrict__ ranges, PixelState ps_all,
                                                          int max_per_pixel, int32_t* __restrict__ out_ids_all) {
    __shared__ RecordLds L;
    __shared__ int32_t s_stash[6 * kBatch];
    const uint32_t gsub = global_tile_of_block(nullptr);
    const BatchItem& B = items[item_of_sub(items, nitems, gsub)];
    const ViewUniforms& U = B.U;
    const float* __restrict__ pos = B.pos;
    const PixelState ps = item_pixel_state(ps_all, B);
    int32_t* __restrict__ out_ids = out_ids_all + B.pixel0 * max_per_pixel;
    const SubTile st = sub_tile(B, gsub);
    const int lane = threadIdx.x;
    const int lx = lane & 7, ly = lane >> 3;
    const int x = st.x0 + lx, y = st.y0 + ly;
    const bool valid = x < U.width && y < U.height;
    int32_t* stash = U.tile_shift != 3 ? s_stash : nullptr;
    const uint2 rg = ranges[st.gtile];
    const int pix = valid ? pixel_index(U, x, y) : 0;
    const uint32_t last = valid ? ps.last_rec[pix] : 0u;
    int k = 0;
    for (uint32_t base = rg.x; base < rg.y; base += kBatch) {
        const int cnt = (int)min((uint32_t)kBatch, rg.y - base);
        if (lane < cnt) {
            RecordData d;
            const RecordIn r = load_record(recs, base + lane);
            fetch_record_data(pos, nullptr, r, d);
            if (stash != nullptr)
                store_record<true>(U, r, d, L, lane, stash);
            else
                store_record<false>(U, r, d, L, lane, nullptr);
        }
        __syncthreads();
        if (stash != nullptr) {
            subtile_masks(L, stash, cnt, x, y, valid, lane);
            __syncthreads();
        }
        uint32_t m = valid ? lane_coverage(L, cnt, lane) : 0u;
        while (m) {
            const int j = __ffs(m) - 1;
            m &= m - 1u;
            if (base + j >= last) break;
            if (k < max_per_pixel)
                out_ids[(int64_t)pix * max_per_pixel + k] = (int32_t)recs[2 * (int64_t)(base + j) + 1].x;
            k++;
        }
        __syncthreads();
    }
    if (valid)
        for (; k < max_per_pixel; k++) out_ids[(int64_t)pix * max_per_pixel + k] = -1;
}

// Workgroups of `kernel` the GPU holds at once (the size of a persistent launch), computed once per kernel and
// device.
static unsigned resident_groups(const void* kernel) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, unsigned> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({kernel, dev});
    if (it != cache.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, 0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const unsigned n = (unsigned)(per_cu > 0 ? per_cu : 1) * (unsigned)(cus > 0 ? cus : 1);
    cache[{kernel, dev}] = n;
    return n;
}

// One launch of a tile kernel: a workgroup per tile, or (queue != NULL) a persistent grid whose workgroups claim
// tiles from the queue counter, zeroed here first (a fill kernel, so that the launch can be graph-captured).
template <typename Kernel, typename... Args>
static void launch_tiles(Kernel grid_kernel, Kernel queue_kernel, int64_t ntiles, uint32_t* queue, hipStream_t s,
                         Args... args) {
    if (queue == nullptr) {
        hipLaunchKernelGGL(grid_kernel, dim3((unsigned)ntiles), dim3(64), 0, s, args..., queue, (uint32_t)ntiles);
        return;
    }
    unsigned grid = resident_groups(reinterpret_cast<const void*>(queue_kernel));
    if (grid == 0 || grid > (uint64_t)ntiles) grid = (unsigned)ntiles;
    fill_u32(queue, 0u, 1, s);
    hipLaunchKernelGGL(queue_kernel, dim3(grid), dim3(64), 0, s, args..., queue, (uint32_t)ntiles);
}

template <bool kCell, bool kSub, int kRound>
static void launch_forward_round(const BatchItem* items, int nitems, int64_t ntiles, const uint32_t* order,
                                 const uint4* recs, const uint2* ranges, PixelState ps, uint32_t* queue,
                                 hipStream_t s) {
    launch_tiles(k_render_forward<kCell, kRound, kSub, false>, k_render_forward<kCell, kRound, kSub, true>, ntiles,
                 queue, s, items, nitems, order, recs, ranges, ps);
}

template <bool kCell, bool kSub>
static void launch_forward_kind(const BatchItem* items, int nitems, int64_t ntiles, const uint32_t* order,
                                const uint4* recs, const uint2* ranges, PixelState ps, int round, uint32_t* queue,
                                hipStream_t s) {
    if (round == 64)
        launch_forward_round<kCell, kSub, 64>(items, nitems, ntiles, order, recs, ranges, ps, queue, s);
    else if (round == 256)
        launch_forward_round<kCell, kSub, 256>(items, nitems, ntiles, order, recs, ranges, ps, queue, s);
    else
        launch_forward_round<kCell, kSub, kRoundDefault>(items, nitems, ntiles, order, recs, ranges, ps, queue, s);
}

void launch_render_forward(const BatchItem* items, int nitems, int64_t ntiles, const uint32_t* order,
                           const uint4* recs, const uint2* ranges, PixelState ps, uint32_t colour_kinds,
                           hipStream_t s, int round, bool sub, uint32_t* queue) {
    if (ntiles <= 0) return;
    if (colour_kinds & kColoursPerVertex) {
        if (sub)
            launch_forward_kind<false, true>(items, nitems, ntiles, order, recs, ranges, ps, round, queue, s);
        else
            launch_forward_kind<false, false>(items, nitems, ntiles, order, recs, ranges, ps, round, queue, s);
    }
    if (colour_kinds & kColoursPerCell) {
        if (sub)
            launch_forward_kind<true, true>(items, nitems, ntiles, order, recs, ranges, ps, round, queue, s);
        else
            launch_forward_kind<true, false>(items, nitems, ntiles, order, recs, ranges, ps, round, queue, s);
    }
}

template <bool kRef, bool kCell, bool kSub>
static void launch_adjoint_kind(const BatchItem* items, const AdjointTarget* targets, int nitems, int64_t ntiles,
                                const uint32_t* order, const uint4* recs, const uint2* ranges, PixelState ps,
                                int use_abs, uint32_t* queue, hipStream_t s) {
    launch_tiles(k_render_adjoint<kRef, kCell, kSub, false>, k_render_adjoint<kRef, kCell, kSub, true>, ntiles,
                 queue, s, items, targets, nitems, order, recs, ranges, ps, use_abs);
}

template <bool kSub>
static void launch_adjoint_sub(const BatchItem* items, const AdjointTarget* targets, int nitems, int64_t ntiles,
                               const uint32_t* order, const uint4* recs, const uint2* ranges, PixelState ps,
                               int use_abs, int mode, uint32_t colour_kinds, uint32_t* queue, hipStream_t s) {
    if (colour_kinds & kColoursPerVertex) {
        if (mode == 1)
            launch_adjoint_kind<true, false, kSub>(items, targets, nitems, ntiles, order, recs, ranges, ps, use_abs,
                                                   queue, s);
        else
            launch_adjoint_kind<false, false, kSub>(items, targets, nitems, ntiles, order, recs, ranges, ps, use_abs,
                                                    queue, s);
    }
    if (colour_kinds & kColoursPerCell)  // exact adjoint only (the caller rejects mode 1 for such items)
        launch_adjoint_kind<false, true, kSub>(items, targets, nitems, ntiles, order, recs, ranges, ps, use_abs,
                                               queue, s);
}

void launch_render_adjoint(const BatchItem* items, const AdjointTarget* targets, int nitems, int64_t ntiles,
                           const uint32_t* order, const uint4* recs, const uint2* ranges, PixelState ps, int use_abs,
                           int mode, uint32_t colour_kinds, hipStream_t s, bool sub, uint32_t* queue) {
    if (ntiles <= 0) return;
    if (sub)
        launch_adjoint_sub<true>(items, targets, nitems, ntiles, order, recs, ranges, ps, use_abs, mode,
                                 colour_kinds, queue, s);
    else
        launch_adjoint_sub<false>(items, targets, nitems, ntiles, order, recs, ranges, ps, use_abs, mode,
                                  colour_kinds, queue, s);
}

// grad_pos[v] += acc[v][0..2], grad_col[v] += acc[v][3..6] (grad_col NULL: positions only).  One thread per
// vertex, so plain read-modify-writes.
__global__ __launch_bounds__(256) void k_fold_vertex_grads(const float4* __restrict__ acc, int64_t V,
                                                           float* __restrict__ grad_pos,
                                                           float* __restrict__ grad_col) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    const float4 a = acc[2 * v], b = acc[2 * v + 1];
    grad_pos[3 * v + 0] += a.x;
    grad_pos[3 * v + 1] += a.y;
    grad_pos[3 * v + 2] += a.z;
    if (grad_col != nullptr) {
        grad_col[4 * v + 0] += a.w;
        grad_col[4 * v + 1] += b.x;
        grad_col[4 * v + 2] += b.y;
        grad_col[4 * v + 3] += b.z;
    }
}

void launch_fold_vertex_grads(const float* acc, int64_t V, float* grad_pos, float* grad_col, hipStream_t s) {
    if (V <= 0) return;
    hipLaunchKernelGGL(k_fold_vertex_grads, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(acc), V, grad_pos, grad_col);
}

void launch_export_fragments(const BatchItem* items, int nitems, int64_t ntiles, const uint4* recs,
                             const uint2* ranges, PixelState ps, int max_per_pixel, int32_t* out_ids, hipStream_t s) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_export_fragments, dim3((unsigned)ntiles), dim3(64), 0, s, items, nitems, recs, ranges,
                       ps, max_per_pixel, out_ids);
}

}  // namespace dtv
