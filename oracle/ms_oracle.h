/*
 * ms_oracle.h -- CPU restatement of MorphStore's compression-aware path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against: only tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py may load it.  The product (engine_amd) never
 * links or calls it.
 *
 * Every function restates the reference algorithm named in its comment
 * (paths relative to the reference root).  Arithmetic is uint64_t mod 2^64.
 * Buffers are host memory; images are relative to a 64-byte aligned base.
 */
#ifndef MS_ORACLE_H
#define MS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/morphstore_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A host column: byte image + column_meta_data (column.h:76-109). */
typedef struct oc_col {
    uint8_t *data;        /* calloc'd, size_alloc bytes */
    uint64_t count;
    uint64_t size_used;
    uint64_t size_compr;
    uint64_t size_alloc;
    ms_format_desc f;
} oc_col;

/* math.h:108-113 */
unsigned oc_effective_bitwidth(uint64_t v);

/* F::get_size_max_byte, get_size_max_byte_any_len (format.h:114-151),
 * F::m_BlockSize, create_data_uncompr_start (column.h:182-196). */
int oc_size_max_byte(const ms_format_desc *f, uint64_t n, uint64_t *out);
int oc_size_max_byte_any_len(const ms_format_desc *f, uint64_t n, uint64_t *out);
int oc_block_size(const ms_format_desc *f, uint64_t *out);
int oc_rest_offset(const ms_format_desc *f, uint64_t compr_end, uint64_t *out);

/* morph_batch_t<ve, F, uncompr_f> / <ve, uncompr_f, F>.  `out` must be
 * zero-filled.  n must be a multiple of F's block size (rle: any n). */
int oc_morph_batch_compress(const ms_format_desc *f, const uint64_t *in,
                            uint64_t n, uint8_t *out, uint64_t *out_bytes);
int oc_morph_batch_decompress(const ms_format_desc *f, const uint8_t *in,
                              uint64_t n, uint64_t *out, uint64_t *in_bytes);

/* Column level (morph.h:157-241, rle.h:155-297). */
int oc_col_from_values(oc_col *out, const uint64_t *v, uint64_t n);
int oc_col_from_image(oc_col *out, const ms_format_desc *f, const uint8_t *image, uint64_t count,
                      uint64_t size_used, uint64_t size_compr);
int oc_morph_compress(oc_col *out, const ms_format_desc *dst,
                      const uint64_t *in, uint64_t n);
int oc_morph_decompress(const oc_col *in, uint64_t *out);
/* blockwise_individual_f (blockwise_individual.h:113-600): the default
 * format of a format code (MS_ERR_INVALID for codes without one), the
 * column-level morph with given codes, the codes of a column, and the
 * format's selector over one block. */
int oc_bwi_inner(unsigned code, ms_format_desc *in);
int oc_morph_compress_codes(oc_col *out, const ms_format_desc *dst, const uint64_t *in, uint64_t n,
                            const uint8_t *codes);
int oc_bwi_codes(const oc_col *c, uint8_t *codes, uint64_t cap, uint64_t *nblocks);
int oc_bwi_choose(const ms_format_desc *f, const uint64_t *in, unsigned *code);
uint64_t oc_count_values_uncompr(const oc_col *c);
/* column<F>::prepare_for_random_access (column.h:241-268); *acted = 1 if
 * the column changed. */
int oc_prepare_for_random_access(oc_col *c, int *acted);
void oc_col_free(oc_col *c);

/* Write-iterator output of m values in format f (write_iterator.h:66-306,
 * delta.h:369-490, uncompr.h:125-221).  Appends nothing after done();
 * *append_off receives done()'s second tuple element. */
int oc_wit_emit(oc_col *out, const ms_format_desc *f, const uint64_t *v,
                uint64_t m, uint64_t extra_capacity, uint64_t *append_off);

/* Operators. */
int oc_select(oc_col *out, uint32_t cmp, uint64_t val,
              const ms_format_desc *out_f, const oc_col *in, uint32_t lanes);
int oc_between(oc_col *out, uint32_t cmp_lo, uint64_t val_lo, uint32_t cmp_hi, uint64_t val_hi,
               const ms_format_desc *out_f, const oc_col *in, uint32_t lanes);
int oc_project(oc_col *out, const ms_format_desc *out_f, const oc_col *data,
               const oc_col *pos, uint32_t lanes);
int oc_agg_sum(oc_col *out, const oc_col *in);
int oc_intersect_sorted(oc_col *out, const ms_format_desc *out_f,
                        const oc_col *l, const oc_col *r);
int oc_merge_sorted(oc_col *out, const ms_format_desc *out_f,
                    const oc_col *l, const oc_col *r);
/* scalar/group_uncompr.h:44-113 (in_gr NULL: unary group), scalar/
 * join_uncompr.h:44-128 (nested loops; semi join); uncompressed outputs. */
int oc_group(oc_col *out_gr, oc_col *out_ext, const oc_col *in_gr, const oc_col *in_data);
int oc_join(oc_col *out_l, oc_col *out_r, const oc_col *l, const oc_col *r);
int oc_semi_join(oc_col *out, const oc_col *l, const oc_col *r);
int oc_calc_binary(oc_col *out, uint32_t op, const ms_format_desc *out_f,
                   const oc_col *l, const oc_col *r);

/* Fused select + agg_sum by decompression, and a block-streaming version
 * for dynamic_vbp / static_vbp / uncompr columns that never materialises
 * the column (the CPU baseline). */
int oc_select_sum(uint64_t *result, uint32_t cmp, uint64_t val, const oc_col *in);
int oc_select_sum_stream(uint64_t *result, uint32_t cmp, uint64_t val,
                         const oc_col *in, int threads);

/* The CPU baseline (ms_cpu.c): streaming select + sum over dynamic_vbp,
 * delta+dynamic_vbp, for+dynamic_vbp (8 lanes), rle and uncompr columns
 * with `threads` threads; simd 0 scalar, 1 AVX-512 (MS_ERR_INVALID where
 * the CPU lacks it or the layout is not streamed). */
int oc_cpu_has_avx512(void);
int oc_cpu_select_sum(uint64_t *result, uint32_t cmp, uint64_t val, const oc_col *in,
                      int threads, int simd);

#ifdef __cplusplus
}
#endif

#endif /* MS_ORACLE_H */
