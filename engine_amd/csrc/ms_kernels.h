// ms_kernels.h -- launch entry points of the kernels (host side), used by
// the C-ABI layer.  Every launcher is stream ordered and reads the counts
// it needs from the columns' device meta data; host-side bounds only size
// the grids.
#pragma once

#include <vector>

#include "ms_internal.h"
#include "ms_device.hpp"

namespace ms {

struct temps;   // stream-ordered temporaries of one operator call

// Thin RAII-free arena: every temporary of a call is freed at the end of
// the call on the call's stream (hipFreeAsync keeps the ordering).
struct temps {
    hipStream_t s;
    void *ptrs[192];
    int n = 0;
    explicit temps(hipStream_t st) : s(st) {}
    void *get(size_t bytes);
    ~temps();
};

int colv_of(const ms_col *c, ms_colv &v);

// Plans: true while the calling thread captures operators into a hipGraph.
bool plan_capturing();
// While capturing: the plan calls release(obj) when it is freed, so an
// object whose device memory the captured graph reads can refuse to change
// or free that memory until then.
int plan_hold(void (*release)(void *), void *obj);

// device memory of the operators (from a recording plan's arena if any)
void *dev_alloc(size_t bytes, hipStream_t s);
void dev_free(void *p, hipStream_t s);

// decode side -------------------------------------------------------
int launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t s);
// Caller-provided device buffers of launch_select_sum: `part` holds
// SS_MAX_PARTS per-workgroup partial sums; `tmp` (tmp_values values) holds
// the decoded deltas of a delta_f column without a unit decoder.  Every
// kernel writes what it later reads, so neither needs clearing.
constexpr uint64_t SS_MAX_PARTS = 8192;
struct ss_buffers {
    uint64_t *part;
    uint64_t *tmp;
    uint64_t tmp_values;
};
// *d_result (+)= sum of x_i with pred(x_i); the result is zeroed first
// unless `accumulate`.  With `bufs` no device memory is allocated.
int launch_select_sum(const ms_col *in, const ms_pred &p, uint64_t *d_result, bool accumulate, hipStream_t s,
                      const ss_buffers *bufs = nullptr);
// writes the column's `count` values to d_out (uncompressed); with `om`
// also the whole meta of the uncompr_f column d_out belongs to
int launch_decompress(const ms_col *in, uint64_t *d_out, hipStream_t s, ms_dev_meta *om = nullptr);
// ascending positions with pred into d_pos (capacity >= count bound);
// *d_cnt receives their number
int launch_select_positions(const ms_col *in, const ms_pred &p, uint64_t *d_pos, uint64_t *d_cnt, temps &sc);
// The selected rows of a select as its count pass leaves them, without
// materialising positions: 8 ballots per 512-row unit (bit 8 gl + j of
// ballot m = row 64 gl + 8 m + j of the unit) and uoff[u] = the number of
// selected rows before unit u (uoff[*d_units] = all of them).  Ranks are
// positions in the select's ascending output.
struct sel_flags {
    const uint64_t *flags;
    const uint64_t *uoff;
    const uint64_t *d_units;
    uint64_t units_bound;    // host bound of *d_units
};
// true when launch_select_flags takes `in` (8-lane dynamic_vbp / for_f)
bool select_flags_ok(const ms_col *in);
int launch_select_flags(const ms_col *in, const ms_pred &p, sel_flags &sf, temps &sc);
// rebuilds a delta_f column's tile carries from its image and index
int launch_delta_carries(const ms_col *in, temps &sc);
// exclusive prefix sums of n u64 values (in-place allowed)
int launch_exclusive_scan(const uint64_t *d_in, uint64_t *d_out, uint64_t n, temps &sc);
// the same over the first min(*d_n + extra, bound) values (device count)
int launch_exclusive_scan_n(const uint64_t *d_in, uint64_t *d_out, const uint64_t *d_n, uint64_t extra,
                            uint64_t bound, temps &sc);

// encode side -------------------------------------------------------
// Emits the m values d_v[0 .. *d_m) into `out` (already allocated, format
// out->f) as `chunk`-value morph_batch calls (chunk 0: one call for the
// whole compressed part, i.e. a column-level morph) followed by the raw
// rest, and sets out's meta.  *d_append receives the write-iterator's
// append offset (done()'s second element).
int launch_emit(ms_col *out, const uint64_t *d_v, const uint64_t *d_m, uint64_t m_bound, uint64_t chunk,
                uint64_t *d_append, temps &sc);
// The select's split of its input (select_compr.h:209-250): rows below
// the scalar remainder's first row go through the write-iterator, the
// positions of the remainder's selected rows are appended raw behind it.
struct sel_split {
    const ms_dev_meta *in_meta;
    ms_geom in_g;
    uint32_t lanes;
    uint64_t *d_cnt, *d_na, *d_flag;   // out: selected rows, those written by the iterator, remainder present
};
// The select's output from its flags (the positions of its selected rows),
// for the formats emit_flags_ok takes (dynamic_vbp_f and delta_f / for_f
// over it): the split, the write-iterator's emission of ranks [0, *d_na)
// and the raw append of ranks [*d_na, *d_cnt).  pos (capacity >= m_bound)
// receives only the positions other kernels read by rank: the delta tile
// predecessors and the ranks behind the compressed part.
bool emit_flags_ok(const ms_format_desc &f);
int launch_emit_flags(ms_col *out, const sel_flags &sf, const sel_split &sp, uint64_t *pos, uint64_t m_bound,
                      uint64_t chunk, temps &sc);
// rle_f compression of d_v[0 .. *d_m)
int launch_rle_compress(ms_col *out, const uint64_t *d_v, const uint64_t *d_m, uint64_t m_bound, temps &sc);

// One morph_batch call of format f (dynamic_vbp_f, or delta_f / for_f over
// it) per chunk of K values; chunk c is v[cmap[c] * K .. + K) and its
// delta predecessors start at zero.  launch_chunk_sizes fills the state
// (temporaries of sc) and each chunk's bytes into d_csz (may be null);
// launch_chunk_pack then writes chunk c at image offset d_cout[c], its
// block index entries (absolute offsets) to index_out and for_f
// references to refs_out, both in chunk order.
struct chunk_emit {
    ms_format_desc f;
    const uint64_t *v;
    const uint32_t *cmap;
    uint64_t nchunks, K;
    uint64_t *d_m, *sz, *start, *refs;
    uint32_t *bws;   // bit width per innermost block
    uint64_t nb;     // innermost blocks
};
int launch_chunk_sizes(chunk_emit &ce, uint64_t *d_csz, temps &sc);
int launch_chunk_pack(const chunk_emit &ce, const uint64_t *d_cout, uint8_t *out, uint64_t cap, uint64_t *status,
                      uint64_t *index_out, uint64_t *refs_out, hipStream_t s);

// blockwise_individual_f (ms_bwi.hip) ------------------------------
// The default format of a format code; false for codes without one.
bool bwi_inner(uint32_t code, ms_format_desc &d);
bool bwi_code_ok(uint32_t code);
// The device view table of a column (inside its allocation, c->vmeta on):
// meta of every view and of the 3 groups, then payload offsets, sizes,
// the per-kind lists of block numbers and the codes.
struct bwi_dev_tab {
    ms_dev_meta *meta;
    uint64_t *q, *psz;
    uint32_t *lists;
    uint8_t *codes;
};
uint64_t bwi_tab_bytes(uint64_t nb);
bwi_dev_tab bwi_tab(const ms_col *c);
uint64_t bwi_index_entries(const ms_bwi &h);
uint64_t bwi_refs_entries(const ms_bwi &h);
// Compression in two steps around the allocation of the output: bwi_plan
// fixes the codes (given, or chosen by the format's selector), the views'
// sizes and offsets (host copies in w.h); bwi_pack writes the column.
struct bwi_work {
    const uint64_t *v;
    ms_bwi h;
    std::vector<uint8_t> codes;
    std::vector<uint32_t> lists;
    uint8_t *d_codes;
    uint32_t *d_lists;
    uint64_t *d_psz, *d_q;
    chunk_emit ce[3];
    uint8_t *gbw;
    uint64_t *gs_cs;
    uint64_t base0;   // image offset of the first code byte
};
int bwi_plan(const ms_format_desc &f, const uint64_t *d_v, uint64_t n, const uint8_t *h_codes, uint64_t base0,
             bwi_work &w, temps &sc);
int bwi_pack(bwi_work &w, ms_col *c, temps &sc);
// Only the image bytes [base0, w.h.C) (a preparation for random access).
int bwi_write_image(bwi_work &w, uint8_t *img, uint64_t cap, uint64_t *status, uint64_t *index_out,
                    uint64_t *refs_out, temps &sc);
// Upload of a host image: the walk (view_walk: the block index walk of one
// view's format over `count` values in `avail` bytes, reporting its bytes),
// then the device table and the delta carries.
struct bwi_host {
    ms_bwi h;
    std::vector<uint8_t> codes;
    std::vector<uint32_t> lists;
    std::vector<uint64_t> idx, refs;
};
int bwi_walk_image(const ms_format_desc &f, const uint8_t *img, uint64_t count, uint64_t size_compr, bwi_host &out,
                   int (*view_walk)(const ms_format_desc &, const uint8_t *, uint64_t, uint64_t,
                                    std::vector<uint64_t> &, std::vector<uint64_t> &, uint64_t *));
int bwi_finish_upload(ms_col *c, bwi_host &hb, temps &sc);
int launch_bwi_decompress(const ms_col *in, uint64_t *d_out, hipStream_t s);
int launch_bwi_select_sum(const ms_col *in, const ms_pred &p, uint64_t *d_result, bool accumulate, hipStream_t s);

// static_vbp<vbp_padding_l>, k_wise_ns, group_simple (ms_xfmt.hip): the
// column-level morph of d_v[0 .. *d_m) into `out` (the call starts base0
// bytes past the image base), and decompression of the compressed part
int launch_xfmt_compress(ms_col *out, const uint64_t *d_v, const uint64_t *d_m, uint64_t m_bound, uint64_t base0,
                         temps &sc);
int launch_xfmt_decompress(const ms_col *in, uint64_t *d_out, hipStream_t s);

// operators ---------------------------------------------------------
int launch_gather(const ms_col *data, const uint64_t *d_pos, const uint64_t *d_np, uint64_t np_bound,
                  uint64_t *d_out, uint64_t *d_status, temps &sc);
int launch_intersect(const uint64_t *a, const uint64_t *d_na, uint64_t na_bound, const uint64_t *b,
                     const uint64_t *d_nb, uint64_t nb_bound, uint64_t *d_out, uint64_t *d_cnt, temps &sc);
int launch_merge(const uint64_t *a, const uint64_t *d_na, uint64_t na_bound, const uint64_t *b,
                 const uint64_t *d_nb, uint64_t nb_bound, uint64_t *d_out, uint64_t *d_cnt, temps &sc);
int launch_calc(uint32_t op, const uint64_t *a, const uint64_t *b, const uint64_t *d_n, uint64_t n_bound,
                uint64_t *d_out, hipStream_t s);
// group ids in order of first appearance and the groups' first positions
// of n values (gr: the ids of a previous grouping, or null)
int launch_group(const uint64_t *data, const uint64_t *gr, uint64_t n, uint64_t *out_gr, uint64_t *out_ext,
                 uint64_t *d_ngroups, temps &sc);
// equal pairs of outer x inner positions in nested-loop order (semi: only
// the outer positions with a match, into *out_outer); synchronises
int launch_join(const uint64_t *outer, uint64_t na, const uint64_t *inner, uint64_t nb, bool semi,
                uint64_t **out_outer, uint64_t **out_inner, uint64_t *d_total, uint64_t *total, temps &sc);
// small device-side bookkeeping kernels
int launch_copy_u64(uint64_t *dst, const uint64_t *src, hipStream_t s);
int launch_set_meta(ms_dev_meta *m, const uint64_t *d_count, uint64_t size_per_value, uint64_t size_compr_is_used,
                    hipStream_t s);

}  // namespace ms

int ms_fail(int code, const char *fmt, ...);
int ms_need_init();   // MS_OK once ms_init succeeded

#define MS_HIP(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return ms_fail(MS_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

#define MS_TRY(x)                   \
    do {                            \
        int r_ = (x);               \
        if (r_ != MS_OK) return r_; \
    } while (0)
