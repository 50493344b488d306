// multi_host.hpp -- host side of the multi-job kernels, shared by batches
// (batch.cpp) and the request server (server.cpp): per-job device state,
// the cache of direction/basis tables, round buffers and round tables.
#pragma once

#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "vlb_internal.hpp"

namespace vlb {

// Direction/basis tables shared by every job with the same direction set,
// coefficient count, constants, projection form and slice length.
class RayCache {
public:
    // `built` (may be null) is set when the table was made by this call
    // (its upload is on the context's stream)
    std::shared_ptr<RayTable> get(Context& ctx, const vlb_bake_params& p, int K, uint32_t RS,
                                  bool* built = nullptr);
    size_t bytes() const;

private:
    using Key = std::tuple<int, uint32_t, uint32_t, int, int, int, uint32_t>;
    std::map<Key, std::shared_ptr<RayTable>> tabs_;
};

// Device-resident state of one job: probes, partial sums and outputs.
struct JobState {
    uint32_t scene_id = 0;
    vlb_bake_params params{};
    std::vector<vlb_light> lights;
    int K = 9;
    int npg = 0;              // probe groups
    int ntiles = 0;           // projection tiles = npg * nslices
    int nch = 0;              // trace tiles per projection tile (RS / 32)
    size_t rays = 0;          // P * R
    std::shared_ptr<RayTable> rt;
    ProbeSet ps;
    DevBuf partials, out, out64;
    JobDesc desc{};
    size_t state_bytes() const { return ps.probes.bytes + ps.probe_perm.bytes + partials.bytes + out.bytes + out64.bytes; }
};

// Rays per slice of a multi-job bake: the caller's value, else the whole
// direction set up to 16384 rays (one slice: the fewest partials).
uint32_t job_slice_rays(const vlb_bake_params& p, uint32_t R);
// Device bytes a job will hold (probes, partials, outputs), for admission.
size_t job_state_bytes(const vlb_bake_params& p, bool out64);
// Validates p, uploads the job's probes and allocates its buffers.
void init_job(Context& ctx, RayCache& cache, uint32_t scene, const vlb_bake_params& p, bool out64,
              JobState& js);

// A round's tables in host form.
struct RoundTable {
    WorkHeader hdr{};
    std::vector<Segment> seg;
    std::vector<MergeItem> merge;
    size_t rays = 0;          // primary rays of the round
};
// Appends projection tiles [pt0, pt0 + npt) of job `job` to a round.
void round_add(RoundTable& r, int job, const JobState& js, int pt0, int npt);
// Adds a job to the round's merge list.
void round_merge(RoundTable& r, int job, const JobState& js);
// Serialises a round table (header, segments, merge items) for upload;
// returns the byte offsets of the three parts.
size_t round_bytes(const RoundTable& r);
void round_pack(const RoundTable& r, uint8_t* dst, size_t off[3]);
// Fixed layout for at most seg_cap segments and merge_cap merge items
// (round_bytes_cap bytes), so one captured graph serves every round.
size_t round_bytes_cap(size_t seg_cap, size_t merge_cap);
void round_pack_cap(const RoundTable& r, uint8_t* dst, size_t seg_cap, size_t off[3]);

// Buffers of one round in flight: radiance, hit records, visibility, work
// counters, traversal spill and the kernel launch shapes.
struct RoundBufs {
    size_t cap_tiles = 0;     // trace tiles
    int nvis = 0;             // visibility bytes per ray
    uint32_t depth = 0;       // deepest BVH served
    int grid = 0, shade_grid = 0, pgrid = 0, mgrid = 0;
    int fgrid[6] = {};          // fused kernel grids at occupancy 0, 6, 8 (+3: far-first shadows)
    DevBuf rad, hits, vis, wctr, spill, counters;
    void alloc(Context& ctx, size_t cap_tiles, int nvis, uint32_t depth);
    // kernel arguments for the given device tables; focc: -1 for the
    // wavefront kernels, else the fused kernel's occupancy (0, 6, 8); far:
    // the fused kernel's shadow queries far first
    MultiArgs args(const JobDesc* jobs, const uint8_t* table, const size_t off[3], int map,
                   int refill, int chunk, int focc, bool far) const;
};

// trace_variant -> lane map, refill threshold, fused occupancy and shadow
// order: 1..12 the fused kernel (1 + map + 4 * occupancy index, occupancy
// {0, 6, 8}; focc = that occupancy), 13..24 the wavefront kernels (13 + map +
// 4 * refill index, refill {16, 32, 8}; focc = -1), 25..36 as 1..12 with
// far-first shadow queries (far = true).
void multi_variant(uint32_t v, int& map, int& refill, int& focc, bool& far);

}  // namespace vlb
