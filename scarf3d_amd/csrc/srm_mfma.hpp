// Structured-grid SRM as a real GEMM on the MFMA units.
//
// On a grid every harmonic factors per axis:
//   amp_n cos(tx + ty + tz) = Re( e^{i tx} * [amp_n e^{i tz}] * e^{i ty} )
// so with X[n,i] = e^{i tx[n,i]}, Y[n,j] = e^{i ty[n,j]}, Z[n,k] = amp_n e^{i tz[n,k]}
//   F[(j,k), i] = sum_n  Re(Z Y)[n,(j,k)] cos tx[n,i] - Im(Z Y)[n,(j,k)] sin tx[n,i]
// i.e. a GEMM with inner dimension 2N between a generated operand G = Z*Y
// (one row per grid column (j,k)) and the x-table X (one column per x).
// Both operands are split into 16-bit parts whose partial products are
// accumulated in FP32 (hi*hi + hi*lo + lo*hi [+ ...]) and promoted into a
// second FP32 accumulator every 128 harmonics.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "srm_device.hpp"

namespace scarf {

struct MfmaPlan {
    int kernel = 0;
    int parts = 0;          // 16-bit parts per operand value
    int np_ = 0;            // padded harmonic count
    int stages = 0;         // np_ / harmonics per stage
    int nx_pad = 0;         // x extent padded to the block tile
    int tiles_x = 0;        // block tiles along x
    int nbx = 4;            // x tile = 32 * nbx points (64, 96 or 128)
    int tj = 0, tk = 0;     // block column tile: tj values of j by tk values of k
    int tj_shift = 0;       // log2(tj)
    int tiles_j = 0, tiles_k = 0;
    long blocks = 0;
    int nparts = 0;         // statistics partials (one per block)
    size_t x_bytes = 0, y_bytes = 0, z_bytes = 0, table_bytes = 0;
    double zscale = 1.0;    // power of two applied to Z and undone in the scaling (fp16 range)
    // mixed x tiles (ensemble members): x_full tiles of 128 points, then up
    // to two tail tiles of 32 * x_tail[i] points (0: none); tiles_x counts all
    bool mixed = false;
    int x_full = 0;
    int x_tail[2] = {0, 0};
    // first stage the operand tables hold: 0 when they hold every stage; a
    // chunk's start when they hold one chunk of a realization (np_ and the
    // table sizes are then the chunk's capacity)
    int tab_s0 = 0;
};

// Power of two that keeps |Z Y| * 2^12 inside fp16 for the scaled-fp16
// scheme (1 for the bf16 schemes and whenever the largest amplitude is < 8).
double mfma_zscale(int kernel, double amp_max);

// nbx = 2..4 forces the x tile width, tk = 1, 2, 4 or 8 the column tile
// (tj = 256 / tk); other values pick them by shape (least padding)
bool mfma_plan(const Geometry &g, int np_, int kernel, MfmaPlan &plan, std::string &err, int nbx = 0, int tk = 0);
// The same with mixed x tiles: 128-point tiles and at most two tails of 64
// or 96 points, which cover any multiple of 32 points without padding
bool mfma_plan_mixed(const Geometry &g, int np_, int kernel, MfmaPlan &plan, std::string &err, int tk = 0);
// Operand tables of harmonics [h0, h0 + nh) of a record of np_ harmonics per
// field (all of them when nh < 0); h0 and nh are multiples of 16. They are
// stored from stage plan.tab_s0 on.
bool mfma_tables(const MfmaPlan &plan, const Geometry &g, const double *d_harm, int np_, void *d_tab,
                 hipStream_t stream, std::string &err, int h0 = 0, int nh = -1);

// Passes of a field launch over a range of stages (16 harmonics each): a
// realization whose harmonics arrive in chunks runs one launch per chunk.
// Every launch but the first carries in the unscaled FP32 running sums
// stored in the field by the one before; only the last (kFieldFinal)
// scales, rounds and stores the field and writes the statistics partials.
// Chunks start at multiples of 8 stages (the promotion period) and all but
// the last hold a multiple of 8 stages, so each value sums the same products
// in the same order as one launch over every stage: the field is the same
// bit for bit.
constexpr int kFieldCarryIn = 1;
constexpr int kFieldFinal = 2;
constexpr int kFieldChunkHarmonics = 128;  // chunk starts are multiples of this

// Launches tiles first_block .. first_block + max_blocks - 1 (all from
// first_block on when max_blocks is 0) over stages [s0, s0 + nstages) (to
// the end when nstages < 0). Fewer than all tiles leave the field
// incomplete: the tile tuner's timing runs, or a field computed in parts.
template <class Real>
bool mfma_field(const MfmaPlan &plan, const Geometry &g, const void *d_tab, double scaling, Real *d_field,
                double *d_part, hipStream_t stream, std::string &err, long max_blocks = 0, long first_block = 0,
                int s0 = 0, int nstages = -1, int flags = kFieldFinal);

// One realization of an ensemble, prepared on the host.
struct EnsembleMember {
    Geometry g;               // structured grid
    int np_ = 0;              // padded harmonic count (multiple of 64)
    const double *record = nullptr;  // host FP64 record, 5 * np_ values (wx, wy, wz, phase, amp)
    double scaling = 0;
    int rescale = 0;
    double sigma = 0;
    void *d_field = nullptr;  // device output, `Real` elements
    int tk = 0;               // column tile override (0: by shape); x tiles are mixed
    double amp_max = 0;       // largest harmonic amplitude of the record
};

// Tile autotuner. For a grid and an MFMA kernel it times every candidate
// block tile (x width 64/96/128 by column shape tj x tk, tj * tk = 256) on a
// short synthetic harmonic record (the kernels' speed does not depend on
// the values) and keeps the fastest. The tile never changes the field: each
// output value sums the same products in the same order. Choices are cached
// per (kernel, output type, nx, ny, nz) for the process and, when
// SCARF_TUNE_CACHE names a file, across processes. SCARF_TUNE=0 turns tuning
// off (tiles by least padding).
struct TileChoice {
    int nbx = 0, tk = 0;
    double ms = 0;        // time of the timed stages with this tile
    double ms_default = 0;  // the same for the least-padding tile
    bool tuned = false;
};

class TileTuner {
public:
    TileTuner() = default;
    ~TileTuner();
    TileTuner(const TileTuner &) = delete;
    TileTuner &operator=(const TileTuner &) = delete;

    static bool enabled();
    // d_scratch: device room for g.points() values of Real (overwritten)
    template <class Real>
    bool choose(const Geometry &g, int kernel, Real *d_scratch, hipStream_t stream, TileChoice &out, std::string &err);
    // Column tile (out.tk; out.nbx = 0) of a grid that runs inside a full
    // launch with mixed x tiles (an ensemble member): the least padded work
    // per unit of saturated throughput, summed over its x tiles, the
    // throughputs timed once per kernel and output type on a synthetic grid
    // that fills the device with no padding for any candidate.
    template <class Real>
    bool choose_saturated(const Geometry &g, int kernel, hipStream_t stream, TileChoice &out, std::string &err);
    static bool forced(TileChoice &out);  // SCARF_TILE, when set to a valid tile
    // the choice made earlier for this grid (in this process, or in the
    // SCARF_TUNE_CACHE file), without timing anything
    bool cached(const Geometry &g, int kernel, size_t real_bytes, TileChoice &out);
    int tuned_shapes() const { return static_cast<int>(cache_.size()); }

private:
    using Key = std::array<long, 5>;
    void load_disk();
    void save_disk(const Key &k, const TileChoice &c);
    // the synthetic record of np_ harmonics in *rec (made on first use; its
    // SoA layout depends on np_, so each timing length has its own) and the
    // timing events
    bool prepare(double **rec, int np_, hipStream_t stream, std::string &err);
    std::map<Key, TileChoice> cache_;
    bool disk_loaded_ = false;
    double *d_rec_ = nullptr;      // single-field candidates
    double *d_rec_sat_ = nullptr;  // saturated-rate table
    void *d_tab_ = nullptr, *d_part_ = nullptr;
    size_t tab_cap_ = 0, part_cap_ = 0;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    // saturated throughput (point-harmonics per ms) per (kernel, sizeof(Real), nbx, tk)
    std::map<std::array<long, 4>, double> rate_;
    void *d_sat_ = nullptr;  // output of the saturating grid
    size_t sat_cap_ = 0;
};

// Monte Carlo ensembles: operand tables of every member (one launch), one persistent
// grouped-GEMM launch over all (member, tile) work items (sorted by cost,
// largest first, and dealt to the blocks round-robin), per-member
// statistics and rescale. The launch sequence is captured once per ensemble shape into a
// hipGraph and replayed; only the harmonic records change between runs.
class Ensemble {
public:
    Ensemble() = default;
    ~Ensemble();
    Ensemble(const Ensemble &) = delete;
    Ensemble &operator=(const Ensemble &) = delete;

    // stats: 6 values per member {count, mean, variance, info3, info4, rescale factor}
    template <class Real>
    bool run(hipStream_t stream, int kernel, const std::vector<EnsembleMember> &members, std::vector<double> &stats,
             double &device_ms, std::string &err);

    int items() const { return nitems_; }
    bool graph_reused() const { return reused_; }

private:
    void release();
    size_t key_ = 0;
    bool reused_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipEvent_t ev_[2] = {nullptr, nullptr};
    void *d_rec_ = nullptr, *d_tab_ = nullptr, *d_part_ = nullptr, *d_desc_ = nullptr, *d_items_ = nullptr;
    void *d_first_ = nullptr;  // block b's items are d_items_[d_first_[b] .. d_first_[b + 1])
    void *d_stat_ = nullptr;
    void *d_targs_ = nullptr;  // per-member operand-table parameters
    size_t rec_bytes_ = 0;
    int nitems_ = 0, nmembers_ = 0;
    std::vector<unsigned char> desc_host_;
};

}  // namespace scarf
