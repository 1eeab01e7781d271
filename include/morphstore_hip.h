/*
 * morphstore_hip.h -- C-ABI of the MI355X (gfx950) implementation of
 * MorphStore's compression-aware execution path.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (MS_OK on success) and never throws.  Device columns are opaque handles
 * (ms_col) that own a device buffer holding the column's byte image, which
 * is identical to the image the reference keeps in host memory, plus a
 * derived block index that is not part of the image.  Column meta data
 * (count, sizeUsed, sizeCompr) lives in device memory, so operators chain
 * without host synchronisation and a chain of operators can be captured in
 * a hipGraph.  Calls are stream ordered: `stream` is a hipStream_t (NULL =
 * the default stream).  Output columns are allocated by the library and
 * released by the caller with ms_col_free.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   column<F>                      include/core/storage/column.h:45-287
 *   morph<ve,dst,src>              include/core/morphing/morph.h:109-116
 *   my_select_wit_t::apply         include/core/operators/general_vectorized/select_compr.h:108-259
 *   my_project_wit_t::apply        include/core/operators/general_vectorized/project_compr.h:80-261
 *   agg_sum<ve,F>                  include/core/operators/interfaces/agg_sum.h:56-65
 *   intersect_sorted<ve,o,l,r>     include/core/operators/interfaces/intersect.h:52-63
 *   merge_sorted<ve,o,l,r>         include/core/operators/interfaces/merge.h:67-82
 *   calc_binary<op,ve,o,l,r>       include/core/operators/interfaces/calc.h:102-119
 *   exhaustive size selector       include/core/format_selection/exhaustive_size_format_selector.h:35-141
 */
#ifndef MORPHSTORE_HIP_H
#define MORPHSTORE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Status codes                                                        */
/* ------------------------------------------------------------------ */
enum ms_status {
    MS_OK = 0,
    MS_ERR_INVALID = 1,   /* invalid argument or unsupported format        */
    MS_ERR_NOMEM = 2,     /* device or host allocation failed              */
    MS_ERR_HIP = 3,       /* HIP runtime error                             */
    MS_ERR_SIZE = 4,      /* operand counts differ (calc_binary)           */
    MS_ERR_CAPACITY = 5,  /* an input exceeds a buffer or index capacity   */
    MS_ERR_NODEVICE = 6,  /* no usable gfx950 device                       */
    MS_ERR_RANGE = 7      /* a position or value is outside its domain     */
};

/* ------------------------------------------------------------------ */
/* Formats                                                             */
/* ------------------------------------------------------------------ */
enum ms_format_kind {
    MS_UNCOMPR = 0,      /* uncompr_f                       format.h:90-96        */
    MS_STATIC_VBP = 1,   /* static_vbp_f<vbp_l<bw,step>>    static_vbp.h:60-68    */
    MS_DYNAMIC_VBP = 2,  /* dynamic_vbp_f<block,page,step>  dynamic_vbp.h:83-159  */
    MS_DELTA = 3,        /* delta_f<block,step,inner>       delta.h:78-157        */
    MS_FOR = 4,          /* for_f<block,page,inner>         for.h:84-164          */
    MS_RLE = 5,          /* rle_f                           rle.h:42-49           */
    MS_STATIC_VBP_PAD = 6, /* static_vbp_f<vbp_padding_l<bw,step>> vbp_padding.h:52-66 */
    MS_K_WISE_NS = 7,    /* k_wise_ns_f<block>              k_wise_ns.h:43-135    */
    MS_GROUP_SIMPLE = 8, /* group_simple_f<step, uint64_t, 8*step> group_simple.h:64-131 */
    MS_BLOCKWISE_INDIVIDUAL = 9 /* blockwise_individual_f<block, selector, ve> blockwise_individual.h:113-158 */
};

/* format_code (format_selector_commons.h:25-103): the format of one block
 * of a blockwise_individual_f column.  Codes 1..64 are static_vbp_f at that
 * bit width.  Codes 68 and 70 name DEFAULT_DELTA_GROUP_SIMPLE_F and
 * DEFAULT_FOR_GROUP_SIMPLE_F, which the reference's default_formats.h does
 * not define, so their layout is unknown: they are refused.  Code 71 is
 * this library's own: rle_f pairs of the block, delimited by their
 * lengths; the reference has no batch-level rle_f and no such code. */
enum ms_format_code {
    MS_BWI_UNCOMPR = 0,
    MS_BWI_DYNAMIC_VBP = 65,          /* dynamic_vbp_f<512, 64, 8>                    */
    MS_BWI_GROUP_SIMPLE = 66,         /* group_simple_f<8, uint64_t, 64>              */
    MS_BWI_DELTA_DYNAMIC_VBP = 67,    /* delta_f<1024, 8, dynamic_vbp_f<512, 64, 8>>  */
    MS_BWI_DELTA_GROUP_SIMPLE = 68,   /* refused                                      */
    MS_BWI_FOR_DYNAMIC_VBP = 69,      /* for_f<1024, 8, dynamic_vbp_f<512, 64, 8>>    */
    MS_BWI_FOR_GROUP_SIMPLE = 70,     /* refused                                      */
    MS_BWI_RLE = 71                   /* rle_f pairs (extension)                      */
};

/* Format selectors of blockwise_individual_f (the `bw` field). */
enum ms_format_selector {
    MS_BWI_SEL_BEST = 0,    /* exhaustive_size_format_selector_helper<true>  (exhaustive_size_format_selector.h:33-141) */
    MS_BWI_SEL_WORST = 1,   /* exhaustive_size_format_selector_helper<false> */
    MS_BWI_SEL_STATIC = 2   /* static_format_selector_helper<page>           (dummy_format_selectors.h:20-39) */
};

/* Meta byte of a non-existent block in the last page of a dynamic_vbp
 * morph_batch call (vbp_commons.h:54). */
#define MS_VBP_BW_NOBLOCK 0xffu

/*
 * A format descriptor is the run-time form of the reference's format
 * template arguments.
 *   blockwise_individual_f<block, selector, ve> (MS_BLOCKWISE_INDIVIDUAL):
 *       block (a multiple of 1024, at most 2^20), bw = the selector
 *       (ms_format_selector), page = the code of MS_BWI_SEL_STATIC, step = 8
 *       (the 8-lane processing style whose default formats the codes name;
 *       blocks are aligned to its 64-byte vectors)
 *   static_vbp_f<vbp_l<bw,step>>     : bw, step
 *   static_vbp_f<vbp_padding_l<bw,step>> (MS_STATIC_VBP_PAD): bw, step; a
 *       block is 64/bw codes per lane, one 64-bit word per lane
 *   k_wise_ns_f<block>               : block (2: the reference implements
 *       only the SSE variant, 2-wise null suppression)
 *   group_simple_f<step, uint64_t, 8*step> : step (values per group; the
 *       data area is aligned to 8*step bytes from the image base)
 *   dynamic_vbp_f<block,page,step>   : block, page, step.  `step` is the
 *       number of interleaved lanes the payload is packed with.  The
 *       reference ignores t_Step and packs with the vector extension's
 *       element count (dynamic_vbp.h:220); both agree when the format is
 *       used with its matching extension, as all DEFAULT_* formats are.
 *   delta_f<block,step,inner>        : block, step, inner_*
 *   for_f<block,page,inner>          : block, page, inner_*
 *   The inner format of a cascade is described by inner_kind (MS_UNCOMPR,
 *   MS_STATIC_VBP or MS_DYNAMIC_VBP) and inner_bw/inner_block/inner_page/
 *   inner_step with the meaning above.
 */
typedef struct ms_format_desc {
    uint32_t kind;
    uint32_t bw;
    uint32_t block;
    uint32_t page;
    uint32_t step;
    uint32_t inner_kind;
    uint32_t inner_bw;
    uint32_t inner_block;
    uint32_t inner_page;
    uint32_t inner_step;
} ms_format_desc;

/* column_meta_data (column_helper.h:32) */
typedef struct ms_col_meta {
    uint64_t count;       /* m_CountLogicalValues */
    uint64_t size_used;   /* m_SizeUsedByte       */
    uint64_t size_compr;  /* m_SizeComprByte      */
    uint64_t size_alloc;  /* m_SizeAllocByte      */
} ms_col_meta;

typedef struct ms_col ms_col;

/* Comparison operators of select (std::equal_to, std::less, ... as used by
 * select_t / my_select_wit_t; compares are unsigned 64-bit). */
enum ms_cmp {
    MS_CMP_EQ = 0, MS_CMP_NE = 1, MS_CMP_LT = 2,
    MS_CMP_LE = 3, MS_CMP_GT = 4, MS_CMP_GE = 5
};

/* Binary operators of calc_binary (calc.h:102-119), arithmetic mod 2^64.
 * DIV and MOD by zero yield 0 (the reference leaves them undefined). */
enum ms_calc_op {
    MS_OP_ADD = 0, MS_OP_SUB = 1, MS_OP_MUL = 2, MS_OP_DIV = 3,
    MS_OP_MOD = 4, MS_OP_MIN = 5, MS_OP_MAX = 6,
    MS_OP_AND = 7, MS_OP_OR = 8, MS_OP_XOR = 9
};

/* ------------------------------------------------------------------ */
/* Library and column management                                       */
/* ------------------------------------------------------------------ */

/* Selects the HIP device, checks it is gfx950 and sets up the stream-ordered
 * allocation pool.  Must be called once per process before anything else. */
int ms_init(int device);

/* Text of the last error on the calling thread ("" if none). */
const char *ms_last_error(void);

/* Library version string. */
const char *ms_version(void);

/* get_size_max_byte_any_len<F>(count) (format.h:114-151): the pessimistic
 * allocation size of a column of `count` values in format f. */
int ms_size_max_byte_any_len(const ms_format_desc *f, uint64_t count,
                             uint64_t *out);

/* F::get_size_max_byte(count) for count a multiple of F::m_BlockSize. */
int ms_size_max_byte(const ms_format_desc *f, uint64_t count, uint64_t *out);

/* F::m_BlockSize (format.h:68).  rle_f has none: returns MS_ERR_INVALID. */
int ms_block_size(const ms_format_desc *f, uint64_t *out);

/* Uploads a host byte image with its meta data into a new device column and
 * builds the column's block index.  `image` holds meta->size_used bytes. */
int ms_col_upload(ms_col **out, const ms_format_desc *f, const void *image,
                  const ms_col_meta *meta, void *stream);

/* An uncompr_f column holding a copy of `count` values already in device
 * memory (no reference counterpart: device-resident producers). */
int ms_col_from_device(ms_col **out, const uint64_t *d_values, uint64_t count, void *stream);

/* Uploads `count` uncompressed values as an uncompr_f column
 * (meta = {count, 8*count, 0}), as the reference's generators produce. */
int ms_col_upload_values(ms_col **out, const uint64_t *values, uint64_t count,
                         void *stream);

/* Copies the column's meta data to the host (synchronises `stream`). */
int ms_col_get_meta(const ms_col *c, ms_col_meta *meta, void *stream);

/* Copies the byte image [0, size_used) to `image` (capacity bytes) and the
 * meta data to `meta`; synchronises `stream`.  The gap between the
 * compressed part and the uncompressed rest reads as zero bytes. */
int ms_col_download(const ms_col *c, void *image, uint64_t capacity,
                    ms_col_meta *meta, void *stream);

/* The column's format descriptor. */
int ms_col_format(const ms_col *c, ms_format_desc *f);

/* column<F>::prepare_for_random_access (column.h:241-268): the
 * uncompressed rest, zero-padded to a multiple of the block size, is
 * compressed by one morph_batch call (fresh pages / references / delta
 * carry) to the end of the compressed part, inside the gap; the meta data
 * are unchanged and the downloaded image then shows those bytes.  *acted
 * (may be NULL) = 1 if the column changed: not for uncompr_f, an empty rest
 * or a second call.  rle_f (no block size) is refused.  Synchronises
 * `stream`. */
int ms_col_prepare_for_random_access(ms_col *c, int *acted, void *stream);

/* Device address of the byte image (64-byte aligned). */
const void *ms_col_device_data(const ms_col *c);

/* An upper bound of the column's count that is known without a device
 * round trip (exact for uploaded columns). */
uint64_t ms_col_count_bound(const ms_col *c);

/* binary_io<F>::store (binary_io.h:70-85): writes count and sizeUsed as
 * two little-endian uint64, then the sizeUsed bytes of the image (the gap
 * reads as zeros).  Synchronises `stream`. */
int ms_col_store(const ms_col *c, const char *path, void *stream);

/* binary_io<F>::load (binary_io.h:43-68): a column of format f from such a
 * file.  The file holds no sizeCompr (the reference's loaded column keeps
 * 0); the loader derives it as a column-level morph of `count` values lays
 * the image out (floor(count / block) values compressed, the rest raw
 * behind the gap; rle_f: the whole image) and refuses (MS_ERR_INVALID) an
 * image of another size.  Synchronises `stream`. */
int ms_col_load(ms_col **out, const ms_format_desc *f, const char *path, void *stream);

/* Releases the column, stream ordered on the stream it was created on:
 * that stream must still exist (destroy a stream only after freeing the
 * columns created on it; the null stream always exists). */
int ms_col_free(ms_col *c);

/* ------------------------------------------------------------------ */
/* Morph (morph.h:109-241, rle.h:155-297)                               */
/* ------------------------------------------------------------------ */

/* Column-level morph from in's format to `dst`.  Compression writes the
 * compressed main part followed, at the 64-byte aligned offset behind a gap
 * of 2*dst::get_size_max_byte(dst::m_BlockSize), by the uncompressed rest
 * (morph.h:157-197).  A same-format morph returns a new handle sharing
 * nothing with `in` (the reference returns `in` itself). */
int ms_morph(ms_col **out, const ms_format_desc *dst, const ms_col *in,
             void *stream);

/* morph_batch_t<ve, blockwise_individual_f<...>, uncompr_f>
 * (blockwise_individual.h:185-322) at column level with the format codes
 * given instead of chosen by the selector: codes[i] (ms_format_code) for
 * block i, ncodes = floor(count / block).  `dst` must be a
 * blockwise_individual_f.  The learned format selection feeds its
 * per-block picks through this.  Synchronises `stream`. */
int ms_morph_blockwise(ms_col **out, const ms_format_desc *dst, const ms_col *in, const uint8_t *codes,
                       uint64_t ncodes, void *stream);

/* The format codes of a blockwise_individual_f column, one per block, into
 * codes[0 .. cap); *n = the number of blocks. */
int ms_col_block_codes(const ms_col *c, uint8_t *codes, uint64_t cap, uint64_t *n);

/* ------------------------------------------------------------------ */
/* Operators                                                           */
/* ------------------------------------------------------------------ */

/* my_select_wit_t<cmp, ve, out_f, in_f>::apply(in, val): ascending positions
 * i with cmp(x_i, val), written through the selective write-iterator into
 * out_f.  `lanes` is the element count of the emulated vector extension
 * (1 scalar, 2 sse, 4 avx2, 8 avx512): the input rest's last (rest mod
 * lanes) values are processed "scalar" and their positions appended
 * uncompressed, as select_compr.h:209-250 does. */
int ms_select(ms_col **out, uint32_t cmp, uint64_t val,
              const ms_format_desc *out_f, const ms_col *in, uint32_t lanes,
              void *stream);

/* my_between_wit_t<cmp_lower, cmp_upper, ve, out_f, in_f>::apply(in,
 * val_lower, val_upper) (between_compr.h:107-270, interface between.h):
 * ascending positions i with cmp_lower(x_i, val_lower) and
 * cmp_upper(x_i, val_upper), written and laid out as ms_select does.  The
 * comparisons are unsigned; NE is refused (the two would not form one
 * interval). */
int ms_between(ms_col **out, uint32_t cmp_lower, uint64_t val_lower,
               uint32_t cmp_upper, uint64_t val_upper, const ms_format_desc *out_f,
               const ms_col *in, uint32_t lanes, void *stream);

/* my_project_wit_t<ve, out_f, data_f, pos_f>::apply(data, pos):
 * out_i = data[pos_i].  Any data format is accepted (the reference needs
 * random access, i.e. uncompr_f or static_vbp_f); positions must be below
 * data's count (else MS_ERR_RANGE is reported by a later call that reads
 * the column's status, see ms_col_check). */
int ms_project(ms_col **out, const ms_format_desc *out_f, const ms_col *data,
               const ms_col *pos, uint32_t lanes, void *stream);

/* agg_sum<ve, F>(in): an uncompr_f column holding the sum mod 2^64. */
int ms_agg_sum(ms_col **out, const ms_col *in, void *stream);

/* Fused select + agg_sum: *d_result (device pointer) = sum of x_i mod 2^64
 * over all i with cmp(x_i, val), without materialising positions.  Equal to
 * agg_sum(project(in, select(cmp, val, in))). */
int ms_select_sum(uint64_t *d_result, uint32_t cmp, uint64_t val,
                  const ms_col *in, void *stream);

/* The same as ms_select_sum, allocating nothing: no reference counterpart;
 * for callers that cannot allocate device memory (code under hipGraph
 * stream capture, tight serving loops).  Its temporaries live in a fixed
 * per-device array of 8 MiB + 64 KiB set up by ms_init: 8192 per-workgroup
 * partial sums and room for 2^20 decoded deltas, which only a delta_f
 * column whose inner format has no unit decoder needs (such a column of
 * more than 2^20 values is refused with MS_ERR_CAPACITY).  All calls on
 * one device share that array: calls that overlap on one device (other
 * streams or threads) race and must be serialised by the caller. */
int ms_select_sum_noalloc(uint64_t *d_result, uint32_t cmp, uint64_t val,
                          const ms_col *in, void *stream);

/* intersect_sorted<ve, out_f, l_f, r_f>(l, r): multiset intersection of two
 * ascending columns (min multiplicity, as scalar/intersect_uncompr.h:69-80),
 * written through the selective write-iterator into out_f. */
int ms_intersect_sorted(ms_col **out, const ms_format_desc *out_f,
                        const ms_col *l, const ms_col *r, void *stream);

/* merge_sorted<ve, out_f, l_f, r_f>(l, r): multiset union of two ascending
 * columns (an equal pair is emitted once, scalar/merge_uncompr.h:69-96). */
int ms_merge_sorted(ms_col **out, const ms_format_desc *out_f,
                    const ms_col *l, const ms_col *r, void *stream);

/* group<ve, gr_f, ext_f, [in_gr_f,] data_f>([in_gr,] data) (interfaces/
 * group.h:38-99, scalar/group_uncompr.h:44-113): *out_gr holds a group id
 * per value, numbered in order of first appearance of the value (in_gr
 * NULL: unary group) or of the pair (in_gr_i, data_i) (binary group);
 * *out_ext holds each group's first position.  A binary group of inputs
 * of different counts is MS_ERR_SIZE.  Synchronises `stream`. */
int ms_group(ms_col **out_gr, ms_col **out_ext, const ms_format_desc *gr_f,
             const ms_format_desc *ext_f, const ms_col *in_gr,
             const ms_col *in_data, void *stream);

/* nested_loop_join<ve, l_f, r_f>(l, r) (interfaces/join.h:38-75,
 * scalar/join_uncompr.h:44-96): the positions of every equal pair, in
 * the reference's nested-loop order: the outer loop runs over the longer
 * input (the left one when both are equally long), both ascending.
 * Synchronises `stream` (the output size is read back). */
int ms_join(ms_col **out_l, ms_col **out_r, const ms_format_desc *l_f,
            const ms_format_desc *r_f, const ms_col *l, const ms_col *r,
            void *stream);

/* left_semi_nto1_nested_loop_join<ve, out_f>(l, r) (interfaces/join.h:
 * 77-110, scalar/join_uncompr.h:98-128): ascending left positions whose
 * value occurs in r.  Synchronises `stream`. */
int ms_semi_join(ms_col **out, const ms_format_desc *out_f, const ms_col *l,
                 const ms_col *r, void *stream);

/* calc_binary<op, ve, out_f, l_f, r_f>(l, r): out_i = op(l_i, r_i) mod 2^64.
 * Counts must match (MS_ERR_SIZE otherwise, scalar/calc_uncompr.h:84-88):
 * eagerly the call fails; inside a plan the counts are compared on the
 * device and a mismatch yields an empty column whose ms_col_check reports
 * MS_ERR_SIZE. */
int ms_calc_binary(ms_col **out, uint32_t op, const ms_format_desc *out_f,
                   const ms_col *l, const ms_col *r, void *stream);

/* ------------------------------------------------------------------ */
/* Query plans as hipGraphs (no reference counterpart: a query in the   */
/* reference is a chain of operator calls, select_sum_benchmark.cpp:    */
/* 78-118; here such a chain is recorded once and replayed)            */
/* ------------------------------------------------------------------ */
typedef struct ms_plan ms_plan;

/* A plan with an arena of `arena_bytes` device bytes; with 0 the plan first
 * measures: the operators issued between the first ms_plan_begin/end pair
 * run eagerly and the plan records the device memory they allocate; the
 * next begin/end pair then captures them into a hipGraph. */
int ms_plan_create(ms_plan **out, uint64_t arena_bytes);

/* Starts recording the operators issued on `stream` (a created stream, not
 * the default one) by the calling thread.  While capturing, their output
 * columns and temporaries live in the plan's arena (ms_col_free of such a
 * column releases only the handle) and no call may synchronise (ms_col_*
 * meta/download calls, group / join and blockwise_individual_f compression
 * are not allowed). */
int ms_plan_begin(ms_plan *p, void *stream);

/* Ends recording; after a capturing pass, instantiates the graph. */
int ms_plan_end(ms_plan *p, void *stream);

/* Replays the captured operators on `stream`; the output columns created
 * while capturing hold the results of the latest replay. */
int ms_plan_launch(ms_plan *p, void *stream);

/* Arena size (bytes measured so far in measuring mode). */
uint64_t ms_plan_arena_bytes(const ms_plan *p);

/* Releases the graph and the arena (and with it the plan's columns). */
int ms_plan_free(ms_plan *p);

/* ------------------------------------------------------------------ */
/* Batched select + sum over many columns in one persistent launch (no  */
/* reference counterpart: the reference issues one operator call per    */
/* column and query, select_sum_benchmark.cpp:78-118)                   */
/* ------------------------------------------------------------------ */
typedef struct ms_batch ms_batch;

/* Launch and work-queue shape of a batch; picked by ms_batch_tune. */
typedef struct ms_batch_config {
    uint32_t wg_per_cu;        /* persistent workgroups of 256 threads per CU    */
    uint32_t units_per_item;   /* 512-value units per work item (bit-packed), <= 64 */
    uint32_t tiles_per_item;   /* 2048-value tiles per work item (delta_f), <= 16   */
    uint32_t runs_per_item;    /* runs per work item (rle_f)                     */
    uint32_t values_per_item;  /* values per work item (uncompressed parts)      */
    uint32_t waves_per_simd;   /* register budget of the kernel: 0 (the compiler's),
                                  6 or 8 resident waves per SIMD                 */
    uint32_t units_per_item_for;    /* units per item of for_f layouts (0: units_per_item) */
    uint32_t units_per_item_narrow; /* units per item of static_vbp of <= 8 bits (0: units_per_item) */
    uint32_t unroll;           /* units whose words one wave requests together
                                  before decoding: 4 (0) or 2                   */
} ms_batch_config;

int ms_batch_create(ms_batch **out);

/* Adds SELECT SUM(x) WHERE cmp(x, val) over `in` as entry *entry, the index
 * of its result.  Every format is taken except static_vbp_f over
 * vbp_padding_l, k_wise_ns_f, group_simple_f and blockwise_individual_f
 * columns holding group_simple_f blocks (MS_ERR_INVALID: decompress them
 * first); delta_f layouts other than step 8 over 8-lane dynamic_vbp are
 * read difference by difference.  `in` must stay alive while the batch is
 * used. */
int ms_batch_enqueue(ms_batch *b, const ms_col *in, uint32_t cmp, uint64_t val, uint64_t *entry);

/* d_results[e] = the sum of entry e, for every entry: one persistent kernel
 * walks the device work queue, a second one folds the per-item partial
 * sums.  The first run after an enqueue builds the queue (it reads the
 * columns' meta data and synchronises `stream` once); later runs only
 * launch the two kernels and may be recorded into a plan.  A batch must
 * have run once before a plan records it, and while a plan that recorded
 * it is alive its queue is pinned: ms_batch_enqueue, ms_batch_set_config,
 * ms_batch_tune and ms_batch_free return MS_ERR_INVALID until every such
 * plan is freed. */
int ms_batch_run(ms_batch *b, uint64_t *d_results, void *stream);

/* Autotuner: times `reps` runs per candidate shape (coordinate descent over
 * the fields of ms_batch_config), keeps the fastest and returns it. */
int ms_batch_tune(ms_batch *b, uint64_t *d_results, uint32_t reps, ms_batch_config *best, void *stream);

int ms_batch_get_config(const ms_batch *b, ms_batch_config *cfg);
int ms_batch_set_config(ms_batch *b, const ms_batch_config *cfg);

/* Entries, work items of the current queue and the bytes one run reads
 * (the sum of sizeUsed over the entries). */
int ms_batch_stats(const ms_batch *b, uint64_t *entries, uint64_t *items, uint64_t *bytes);

/* Kernel timing: with slots > 0 every run records HIP events around the
 * persistent kernel alone (a ring of `slots` launches; 0 turns it off). */
int ms_batch_timing(ms_batch *b, uint32_t slots);

/* Durations in ms of the latest min(*n, recorded) timed launches, oldest
 * first; *n receives their number.  Waits for those launches. */
int ms_batch_kernel_times(ms_batch *b, float *ms, uint32_t *n);

int ms_batch_free(ms_batch *b);

/* No counterpart in the reference: the engine's own timing of its fused
 * decode, for the learned format selection's labels.  Runs the batch
 * `reps` times with every work item's wall-clock time added to its view
 * (an entry, or one block of a blockwise_individual_f entry, in enqueue
 * order; the raw rest of such an entry is not timed), writes the results
 * as ms_batch_run does and ns[v] = the mean time of view v's items, summed
 * over the waves that ran them, for v < cap; *nviews = the number of
 * views.  Synchronises `stream`. */
int ms_batch_view_times(ms_batch *b, uint64_t *d_results, uint32_t reps, double *ns, uint64_t cap,
                        uint64_t *nviews, void *stream);

/* ------------------------------------------------------------------ */
/* Learned per-chunk format selection (replaces the calibration-driven  */
/* choice of src/calibration and the exhaustive size search of          */
/* exhaustive_size_format_selector.h:35-141 for chunks of 2^16 values)  */
/* ------------------------------------------------------------------ */
#define MS_CHUNK_VALUES 65536u   /* values per chunk                           */
#define MS_GEO_LEN 1024u         /* entries of a geometric threshold table     */
#define MS_NCAND 5               /* candidate formats                          */
#define MS_NFEAT 24              /* features per chunk                         */
#define MS_MLP_IN 24
#define MS_MLP_HIDDEN 64
#define MS_MLP_OUT 10            /* log2 size and log2 decode time per candidate */
#define MS_MLP_NPARAM (24 * 64 + 64 + 64 * 64 + 64 + 64 * 10 + 10)

/* Candidates: static_vbp<max bit width, 8>, dynamic_vbp<512, 2048, 8>,
 * delta<1024, 8, dynamic_vbp<512, 64, 8>>, for<1024, 8, dynamic_vbp<512,
 * 64, 8>> and rle. */
enum ms_candidate { MS_CAND_STATIC = 0, MS_CAND_DVBP = 1, MS_CAND_DELTA = 2, MS_CAND_FOR = 3, MS_CAND_RLE = 4 };

/* Synthetic chunk kinds of the label generator. */
enum ms_gen_kind {
    MS_GEN_UNIFORM = 0,   /* top a bits of a random word                                    */
    MS_GEN_DVBP = 1,      /* per 512 values a bit width in [a, b], top bit forced once       */
    MS_GEN_SORTED = 2,    /* base + running sum of geometric gaps (table), first gap 0        */
    MS_GEN_FOR = 3,       /* base + top a bits                                               */
    MS_GEN_RLE = 4        /* runs of geometric length (table), values = top b bits           */
};

typedef struct ms_chunk_spec {
    uint32_t kind, a, b, table;   /* table: index into the geometric tables */
    uint64_t seed, base;
} ms_chunk_spec;

/* Integer statistics per chunk (MS_ST_COUNT uint64 each). */
enum ms_stat {
    MS_ST_SUM_BW = 0, MS_ST_MAX_BW, MS_ST_SUM_DBW, MS_ST_MAX_DBW, MS_ST_SUM_FBW, MS_ST_MAX_FBW, MS_ST_RUNS,
    MS_ST_NONDEC, MS_ST_ZEROS, MS_ST_MIN, MS_ST_MAX, MS_ST_SUM_BW2, MS_ST_DLESS, MS_ST_FLESS, MS_ST_HIST,
    MS_ST_RESERVED, MS_ST_COUNT
};

/* Writes nchunks chunks of MS_CHUNK_VALUES values to d_out (device) from
 * host specs; geo_tables (host) holds ntables threshold tables. */
int ms_gen_chunks(uint64_t *d_out, const ms_chunk_spec *specs, uint64_t nchunks,
                  const uint64_t *geo_tables, uint32_t ntables, void *stream);

/* Per chunk of d_values (device, nchunks * MS_CHUNK_VALUES values): the
 * MS_ST_* statistics into d_stats; with a host cost table (MS_NCAND x 65
 * floats: ns per 512-value unit by bit width for static/dvbp/delta/for;
 * rle: [0] fixed ns, [1] ns per run) also decode-time labels (MS_NCAND
 * floats per chunk) into d_time_labels. */
int ms_chunk_stats(const uint64_t *d_values, uint64_t nchunks, uint64_t *d_stats,
                   const float *cost_table, float *d_time_labels, void *stream);

/* From the statistics: MS_NFEAT float features and the exact size in bytes
 * of every candidate (MS_NCAND uint64) per chunk; either output may be
 * NULL. */
int ms_chunk_features(const uint64_t *d_stats, uint64_t nchunks, float *d_features,
                      uint64_t *d_sizes, void *stream);

/* The format descriptor of candidate `cand` (static_bw for MS_CAND_STATIC). */
int ms_candidate_format(uint32_t cand, uint32_t static_bw, ms_format_desc *out);

typedef struct ms_mlp ms_mlp;

/* A network with Glorot-uniform weights drawn from splitmix64(seed). */
int ms_mlp_create(ms_mlp **out, uint64_t seed);
int ms_mlp_get_params(const ms_mlp *m, float *host_params);        /* MS_MLP_NPARAM floats */
int ms_mlp_set_params(ms_mlp *m, const float *host_params);        /* resets Adam state    */

/* Loss (1/n) sum_s sum_o w_o (y_so - t_so)^2 over n samples (d_x: n x 24,
 * d_y: n x 10, device) and its gradient (MS_MLP_NPARAM floats, device). */
int ms_mlp_loss_grad(ms_mlp *m, const float *d_x, const float *d_y, const float *out_weights,
                     uint64_t n, float *d_grad, float *d_loss, void *stream);

/* Adam training (lr, 0.9, 0.999, 1e-8) on shuffled mini-batches of `batch`
 * samples (a multiple of 64); epoch_loss (host, `epochs` floats) receives
 * the mean loss of each epoch. */
int ms_mlp_train(ms_mlp *m, const float *d_x, const float *d_y, const float *out_weights, uint64_t n,
                 uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *epoch_loss, void *stream);

/* Predictions (n x 10, device). */
int ms_mlp_predict(const ms_mlp *m, const float *d_x, uint64_t n, float *d_y, void *stream);
int ms_mlp_free(ms_mlp *m);

typedef struct ms_select_params {
    float size_off, size_scale;   /* predicted log2 bytes = y * size_scale + size_off */
    float time_off, time_scale;   /* predicted log2 ns    = y * time_scale + time_off */
    float size_tol;               /* rule 0: candidates within (1 + size_tol) x the smallest size prediction ... */
    float time_margin;            /* ... replace it when predicted faster by more than this fraction */
    float exact_tol;              /* rule 0: fallback to the exact smallest beyond (1 + exact_tol) x it;
                                     rule 1: the candidates admitted, within (1 + exact_tol) x the exact smallest */
    uint32_t rule;                /* 0: pick by predicted size (exact sizes check the pick);
                                     1: the exact smallest, replaced by the admitted candidate predicted
                                        fastest when that is faster by more than time_margin */
} ms_select_params;

/* Per chunk a candidate (bit 7 set: rule 0's exact-size fallback was taken,
 * or rule 1 kept the exact smallest). */
int ms_select_formats(const float *d_pred, const uint64_t *d_sizes, uint64_t n,
                      const ms_select_params *sp, uint8_t *d_choice, void *stream);

/* Reads the column's device status word: MS_OK, or the first error a kernel
 * recorded while producing it (e.g. MS_ERR_RANGE for an out-of-range
 * position in ms_project).  Synchronises `stream`. */
int ms_col_check(const ms_col *c, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MORPHSTORE_HIP_H */
