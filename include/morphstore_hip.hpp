/*
 * morphstore_hip.hpp -- header-only C++ mirror of MorphStore's template API
 * for the compression-aware path, on top of the C-ABI in morphstore_hip.h.
 *
 * Query code written against the reference keeps its shape:
 *
 *   using ve = vectorlib::hip_gfx950;
 *   auto c  = make_column(values, n);                                   // column<uncompr_f>
 *   auto cc = morph<ve, static_vbp_f<vbp_l<12, 1>>>(c);                 // morph.h:109-116
 *   auto p  = my_select_wit_t<less, ve, uncompr_f,
 *                             static_vbp_f<vbp_l<12, 1>>>::apply(cc, 2048);   // select_compr.h:91-112
 *   auto d  = my_project_wit_t<ve, uncompr_f, static_vbp_f<vbp_l<12, 1>>,
 *                              uncompr_f>::apply(cc, p);                // project_compr.h:69-83
 *   auto s  = agg_sum<ve, uncompr_f>(d);                                // agg_sum.h:56-65
 *
 * Formats are the reference's class templates (format tags only: the bytes
 * live on the device); operators return new columns the caller deletes, and
 * a same-format morph returns its input pointer (morph.h:128-151).  Columns
 * live in device memory; their data pointer is a device address.  Calls are
 * ordered on the calling thread's stream (set_stream; default: the null
 * stream).  C-ABI status codes become std::runtime_error here.
 */
#ifndef MORPHSTORE_HIP_HPP
#define MORPHSTORE_HIP_HPP

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "morphstore_hip.h"

namespace morphstore {

inline void ms_check(int rc)
{
    if (rc != MS_OK) throw std::runtime_error(std::string("morphstore-hip: ") + ms_last_error());
}

inline void *&current_stream()
{
    static thread_local void *s = nullptr;
    return s;
}

/* Stream of the calling thread's subsequent calls (a hipStream_t). */
inline void set_stream(void *stream) { current_stream() = stream; }

namespace vectorlib {

/* Processing style of this backend.  Operator outputs reproduce the byte
 * layout the reference writes with a vector extension of t_lanes 64-bit
 * lanes (8: avx512<v512<uint64_t>>, 4: avx2, 2: sse, 1: scalar). */
template <unsigned t_lanes = 8>
struct hip_gfx950_lanes {
    static constexpr unsigned lanes = t_lanes;
};
using hip_gfx950 = hip_gfx950_lanes<8>;

/* Comparison tags with the reference's shape (vector/primitives/compare.h). */
template <class t_ve, int t_granularity = 64> struct equal { static constexpr uint32_t code = MS_CMP_EQ; };
template <class t_ve, int t_granularity = 64> struct nequal { static constexpr uint32_t code = MS_CMP_NE; };
template <class t_ve, int t_granularity = 64> struct less { static constexpr uint32_t code = MS_CMP_LT; };
template <class t_ve, int t_granularity = 64> struct lessequal { static constexpr uint32_t code = MS_CMP_LE; };
template <class t_ve, int t_granularity = 64> struct greater { static constexpr uint32_t code = MS_CMP_GT; };
template <class t_ve, int t_granularity = 64> struct greaterequal { static constexpr uint32_t code = MS_CMP_GE; };

}  // namespace vectorlib

using vectorlib::equal;
using vectorlib::greater;
using vectorlib::greaterequal;
using vectorlib::less;
using vectorlib::lessequal;
using vectorlib::nequal;

// ------------------------------------------------------------------ formats
struct format {};

struct uncompr_f : format {
    static constexpr size_t m_BlockSize = 1;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_UNCOMPR;
        return d;
    }
};

template <unsigned t_bw, unsigned t_step>
struct vbp_l {
    static constexpr unsigned m_Bw = t_bw;
    static constexpr unsigned m_Step = t_step;
};

template <class t_layout>
struct static_vbp_f;

template <unsigned t_bw, unsigned t_step>
struct static_vbp_f<vbp_l<t_bw, t_step>> : format {
    static_assert(t_bw >= 1 && t_bw <= 64, "static_vbp_f: bit width in [1, 64]");
    static constexpr size_t m_BlockSize = 64 * t_step;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_STATIC_VBP;
        d.bw = t_bw;
        d.step = t_step;
        return d;
    }
};

template <size_t t_BlockSizeLog, size_t t_PageSizeBlocks, unsigned t_Step>
struct dynamic_vbp_f : format {
    static constexpr size_t m_BlockSize = t_BlockSizeLog;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_DYNAMIC_VBP;
        d.block = (uint32_t)t_BlockSizeLog;
        d.page = (uint32_t)t_PageSizeBlocks;
        d.step = t_Step;
        return d;
    }
};

template <class t_inner>
inline ms_format_desc with_inner(ms_format_desc d)
{
    const ms_format_desc in = t_inner::desc();
    d.inner_kind = in.kind;
    d.inner_bw = in.bw;
    d.inner_block = in.block;
    d.inner_page = in.page;
    d.inner_step = in.step;
    return d;
}

template <size_t t_BlockSizeLog, unsigned t_step, class t_inner_f>
struct delta_f : format {
    static constexpr size_t m_BlockSize = t_BlockSizeLog;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_DELTA;
        d.block = (uint32_t)t_BlockSizeLog;
        d.step = t_step;
        return with_inner<t_inner_f>(d);
    }
};

template <size_t t_BlockSizeLog, size_t t_PageSizeBlocks, class t_inner_f>
struct for_f : format {
    static constexpr size_t m_BlockSize = t_BlockSizeLog;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_FOR;
        d.block = (uint32_t)t_BlockSizeLog;
        d.page = (uint32_t)t_PageSizeBlocks;
        return with_inner<t_inner_f>(d);
    }
};

struct rle_f : format {
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_RLE;
        return d;
    }
};

/* vbp_padding_l<bw, step> (vbp_padding.h:45-66): the layout of
 * static_vbp_f<vbp_padding_l<bw, step>>. */
template <unsigned t_bw, unsigned t_step>
struct vbp_padding_l {
    static constexpr unsigned m_Bw = t_bw;
    static constexpr unsigned m_Step = t_step;
};

template <unsigned t_bw, unsigned t_step>
struct static_vbp_f<vbp_padding_l<t_bw, t_step>> : format {
    static_assert(t_bw >= 1 && t_bw <= 64, "vbp_padding_l: bit width in [1, 64]");
    static constexpr size_t m_BlockSize = (64 / t_bw) * t_step;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_STATIC_VBP_PAD;
        d.bw = t_bw;
        d.step = t_step;
        return d;
    }
};

/* k_wise_ns_f<t_BlockSizeLog> (k_wise_ns.h:43-135; implemented for 2). */
template <size_t t_BlockSizeLog>
struct k_wise_ns_f : format {
    static_assert(t_BlockSizeLog == 2, "k_wise_ns_f: the reference implements 2-wise null suppression");
    static constexpr size_t m_BlockSize = t_BlockSizeLog;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_K_WISE_NS;
        d.block = (uint32_t)t_BlockSizeLog;
        return d;
    }
};

/* group_simple_f<t_GroupSizeLog, t_base_t, t_AlignmentBytes>
 * (group_simple.h:64-131) with 64-bit values and vector-sized alignment. */
template <size_t t_GroupSizeLog, typename t_base_t, size_t t_AlignmentBytes>
struct group_simple_f : format {
    static_assert(std::is_same<t_base_t, uint64_t>::value, "group_simple_f: uint64_t values");
    static_assert(t_AlignmentBytes == 8 * t_GroupSizeLog, "group_simple_f: aligned to one vector of the group");
    static constexpr size_t m_BlockSize = t_GroupSizeLog;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_GROUP_SIMPLE;
        d.step = (uint32_t)t_GroupSizeLog;
        return d;
    }
};

/* format_code (format_selector_commons.h:25-103): the format of one block
 * of a blockwise_individual_f column; static_vbp_f at bit width bw is
 * code bw.  defaultDeltaGroupSimple and defaultForGroupSimple name formats
 * the reference never defines: the library refuses them. */
enum class format_code : uint8_t {
    uncompr = 0,
    defaultDynamicVbp = 65,
    defaultGroupSimple,
    defaultDeltaDynamicVbp,
    defaultDeltaGroupSimple,
    defaultForDynamicVbp,
    defaultForGroupSimple,
    defaultStaticVbp_1 = 1, defaultStaticVbp_2 = 2, defaultStaticVbp_3 = 3, defaultStaticVbp_4 = 4,
    defaultStaticVbp_5 = 5, defaultStaticVbp_6 = 6, defaultStaticVbp_7 = 7, defaultStaticVbp_8 = 8,
    defaultStaticVbp_9 = 9, defaultStaticVbp_10 = 10, defaultStaticVbp_11 = 11, defaultStaticVbp_12 = 12,
    defaultStaticVbp_13 = 13, defaultStaticVbp_14 = 14, defaultStaticVbp_15 = 15, defaultStaticVbp_16 = 16,
    defaultStaticVbp_17 = 17, defaultStaticVbp_18 = 18, defaultStaticVbp_19 = 19, defaultStaticVbp_20 = 20,
    defaultStaticVbp_21 = 21, defaultStaticVbp_22 = 22, defaultStaticVbp_23 = 23, defaultStaticVbp_24 = 24,
    defaultStaticVbp_25 = 25, defaultStaticVbp_26 = 26, defaultStaticVbp_27 = 27, defaultStaticVbp_28 = 28,
    defaultStaticVbp_29 = 29, defaultStaticVbp_30 = 30, defaultStaticVbp_31 = 31, defaultStaticVbp_32 = 32,
    defaultStaticVbp_33 = 33, defaultStaticVbp_34 = 34, defaultStaticVbp_35 = 35, defaultStaticVbp_36 = 36,
    defaultStaticVbp_37 = 37, defaultStaticVbp_38 = 38, defaultStaticVbp_39 = 39, defaultStaticVbp_40 = 40,
    defaultStaticVbp_41 = 41, defaultStaticVbp_42 = 42, defaultStaticVbp_43 = 43, defaultStaticVbp_44 = 44,
    defaultStaticVbp_45 = 45, defaultStaticVbp_46 = 46, defaultStaticVbp_47 = 47, defaultStaticVbp_48 = 48,
    defaultStaticVbp_49 = 49, defaultStaticVbp_50 = 50, defaultStaticVbp_51 = 51, defaultStaticVbp_52 = 52,
    defaultStaticVbp_53 = 53, defaultStaticVbp_54 = 54, defaultStaticVbp_55 = 55, defaultStaticVbp_56 = 56,
    defaultStaticVbp_57 = 57, defaultStaticVbp_58 = 58, defaultStaticVbp_59 = 59, defaultStaticVbp_60 = 60,
    defaultStaticVbp_61 = 61, defaultStaticVbp_62 = 62, defaultStaticVbp_63 = 63, defaultStaticVbp_64 = 64
};

/* format_selector<ve> (format_selector_commons.h:121-138) and the
 * selectors of exhaustive_size_format_selector.h:33-141 and
 * dummy_format_selectors.h:20-39.  The selection runs in the library; a
 * selector only names it. */
template <class t_vector_extension>
struct format_selector {
};

template <bool t_ChooseBest>
struct exhaustive_size_format_selector_helper {
    template <class t_vector_extension>
    struct selector : format_selector<t_vector_extension> {
        static constexpr uint32_t ms_selector = t_ChooseBest ? MS_BWI_SEL_BEST : MS_BWI_SEL_WORST;
        static constexpr uint32_t ms_code = 0;
    };
};

template <format_code t_FormatCode>
struct static_format_selector_helper {
    template <class t_vector_extension>
    struct selector : format_selector<t_vector_extension> {
        static constexpr uint32_t ms_selector = MS_BWI_SEL_STATIC;
        static constexpr uint32_t ms_code = (uint32_t)t_FormatCode;
    };
};

/* blockwise_individual_f<t_BlockSizeLog, t_format_selector, ve>
 * (blockwise_individual.h:113-158): a format chosen per block of
 * t_BlockSizeLog values among the 8-lane style's default formats. */
template <size_t t_BlockSizeLog, template <class> class t_format_selector, class t_vector_extension>
struct blockwise_individual_f : format {
    static_assert(std::is_base_of<format_selector<t_vector_extension>, t_format_selector<t_vector_extension>>::value,
                  "blockwise_individual_f: the template parameter t_format_selector must be a subclass of "
                  "format_selector");
    static_assert(t_vector_extension::lanes == 8, "blockwise_individual_f: the 8-lane processing style's formats");
    static_assert(t_BlockSizeLog % 1024 == 0 && t_BlockSizeLog <= (1u << 20),
                  "blockwise_individual_f: the logical block size must be a multiple of each default format's");
    static constexpr size_t m_BlockSize = t_BlockSizeLog;
    static ms_format_desc desc()
    {
        ms_format_desc d = {};
        d.kind = MS_BLOCKWISE_INDIVIDUAL;
        d.block = (uint32_t)t_BlockSizeLog;
        d.bw = t_format_selector<t_vector_extension>::ms_selector;
        d.page = t_format_selector<t_vector_extension>::ms_code;
        d.step = 8;
        return d;
    }
};

// ------------------------------------------------------------------ column
/* column<F> (column.h:45-287): a device column in format F. */
template <class F>
class column {
    static_assert(std::is_base_of<format, F>::value, "column: template parameter F must be a subclass of format");

  public:
    explicit column(ms_col *c) : m_Col(c) {}
    column(column<F> const &) = delete;
    column(column<F> &&) = delete;
    column<F> &operator=(column<F> const &) = delete;
    column<F> &operator=(column<F> &&) = delete;
    ~column() { ms_col_free(m_Col); }

    /* Device address of the byte image. */
    const void *get_data() const { return ms_col_device_data(m_Col); }
    size_t get_count_values() const { return meta().count; }
    size_t get_size_used_byte() const { return meta().size_used; }
    size_t get_size_compr_byte() const { return meta().size_compr; }
    /* Values of the uncompressed rest (column.h:205-214). */
    size_t get_count_values_uncompr() const
    {
        const ms_col_meta m = meta();
        if (m.size_compr == m.size_used) return 0;
        size_t bs = 1;
        const ms_format_desc d = F::desc();
        if (d.kind != MS_RLE) ms_check(ms_block_size(&d, &bs));
        return m.count % bs;
    }

    /* The byte image (sizeUsed bytes) in host memory. */
    std::vector<uint8_t> download() const
    {
        const ms_col_meta m = meta();
        std::vector<uint8_t> img(m.size_used ? m.size_used : 1);
        ms_check(ms_col_download(m_Col, img.data(), img.size(), nullptr, current_stream()));
        img.resize(m.size_used);
        return img;
    }

    ms_col *handle() const { return m_Col; }

    /* column.h:241-268: true if the column changed. */
    template <class t_vector_extension>
    bool prepare_for_random_access() const
    {
        int acted = 0;
        ms_check(ms_col_prepare_for_random_access(m_Col, &acted, current_stream()));
        return acted != 0;
    }

  private:
    ms_col *m_Col;
    ms_col_meta meta() const
    {
        ms_col_meta m;
        ms_check(ms_col_get_meta(m_Col, &m, current_stream()));
        return m;
    }
};

/* column<uncompr_f> from host values (the reference's generators produce
 * such columns, e.g. utils/data_generator.h). */
inline const column<uncompr_f> *make_column(const uint64_t *values, size_t count)
{
    ms_col *c = nullptr;
    ms_check(ms_col_upload_values(&c, values, count, current_stream()));
    return new column<uncompr_f>(c);
}

inline const column<uncompr_f> *make_column(const std::vector<uint64_t> &v) { return make_column(v.data(), v.size()); }

/* A column of format F from its byte image and meta data (host memory). */
template <class F>
const column<F> *upload_column(const void *image, size_t count, size_t size_used, size_t size_compr)
{
    const ms_format_desc d = F::desc();
    ms_col_meta m = {count, size_used, size_compr, size_used};
    ms_col *c = nullptr;
    ms_check(ms_col_upload(&c, &d, image, &m, current_stream()));
    return new column<F>(c);
}

/* binary_io<F> (binary_io.h:38-87). */
template <class F>
class binary_io {
  public:
    static const column<F> *load(const std::string &p_Filename)
    {
        const ms_format_desc d = F::desc();
        ms_col *c = nullptr;
        ms_check(ms_col_load(&c, &d, p_Filename.c_str(), current_stream()));
        return new column<F>(c);
    }
    static void store(const column<F> *p_Column, const std::string &p_FileName)
    {
        ms_check(ms_col_store(p_Column->handle(), p_FileName.c_str(), current_stream()));
    }
};

// ------------------------------------------------------------------ operators
template <class t_ve, class t_dst_f, class t_src_f>
struct morph_t {
    static const column<t_dst_f> *apply(const column<t_src_f> *in)
    {
        if constexpr (std::is_same<t_dst_f, t_src_f>::value) {
            return in;   // morph.h:128-151: the input itself
        } else {
            const ms_format_desc d = t_dst_f::desc();
            ms_col *c = nullptr;
            ms_check(ms_morph(&c, &d, in->handle(), current_stream()));
            return new column<t_dst_f>(c);
        }
    }
};

template <class t_ve, class t_dst_f, class t_src_f>
const column<t_dst_f> *morph(const column<t_src_f> *in)
{
    return morph_t<t_ve, t_dst_f, t_src_f>::apply(in);
}

/* Decompressed values in host memory. */
template <class F>
std::vector<uint64_t> values_of(const column<F> *in)
{
    const column<uncompr_f> *u = morph<vectorlib::hip_gfx950, uncompr_f>(in);
    const std::vector<uint8_t> img = u->download();
    std::vector<uint64_t> v(img.size() / 8);
    for (size_t i = 0; i < v.size(); i++) {
        uint64_t x = 0;
        for (int b = 7; b >= 0; b--) x = (x << 8) | img[8 * i + b];
        v[i] = x;
    }
    if (static_cast<const void *>(u) != static_cast<const void *>(in)) delete u;
    return v;
}

template <template <class, int> class t_compare, class t_ve, class t_out_pos_f, class t_in_data_f>
struct my_select_wit_t {
    static const column<t_out_pos_f> *apply(const column<t_in_data_f> *const inDataCol, const uint64_t val,
                                            const size_t outPosCountEstimate = 0)
    {
        (void)outPosCountEstimate;
        const ms_format_desc d = t_out_pos_f::desc();
        ms_col *c = nullptr;
        ms_check(ms_select(&c, t_compare<t_ve, 64>::code, val, &d, inDataCol->handle(), t_ve::lanes,
                           current_stream()));
        return new column<t_out_pos_f>(c);
    }
};

/* my_between_wit_t (between_compr.h:107-270). */
template <template <class, int> class t_compare_lower, template <class, int> class t_compare_upper, class t_ve,
          class t_out_pos_f, class t_in_data_f>
struct my_between_wit_t {
    static const column<t_out_pos_f> *apply(const column<t_in_data_f> *const inDataCol, const uint64_t val_lower,
                                            const uint64_t val_upper, const size_t outPosCountEstimate = 0)
    {
        (void)outPosCountEstimate;
        const ms_format_desc d = t_out_pos_f::desc();
        ms_col *c = nullptr;
        ms_check(ms_between(&c, t_compare_lower<t_ve, 64>::code, val_lower, t_compare_upper<t_ve, 64>::code,
                            val_upper, &d, inDataCol->handle(), t_ve::lanes, current_stream()));
        return new column<t_out_pos_f>(c);
    }
};

/* group (interfaces/group.h:38-99): unary, then binary. */
template <class t_ve, class t_out_gr_f, class t_out_ext_f, class t_in_data_f>
const std::tuple<const column<t_out_gr_f> *, const column<t_out_ext_f> *>
group(const column<t_in_data_f> *const inDataCol, const size_t outExtCountEstimate = 0)
{
    (void)outExtCountEstimate;
    const ms_format_desc dg = t_out_gr_f::desc(), de = t_out_ext_f::desc();
    ms_col *g = nullptr, *e = nullptr;
    ms_check(ms_group(&g, &e, &dg, &de, nullptr, inDataCol->handle(), current_stream()));
    return std::make_tuple(new column<t_out_gr_f>(g), new column<t_out_ext_f>(e));
}

template <class t_ve, class t_out_gr_f, class t_out_ext_f, class t_in_gr_f, class t_in_data_f>
const std::tuple<const column<t_out_gr_f> *, const column<t_out_ext_f> *>
group(const column<t_in_gr_f> *const inGrCol, const column<t_in_data_f> *const inDataCol,
      const size_t outExtCountEstimate = 0)
{
    (void)outExtCountEstimate;
    const ms_format_desc dg = t_out_gr_f::desc(), de = t_out_ext_f::desc();
    ms_col *g = nullptr, *e = nullptr;
    ms_check(ms_group(&g, &e, &dg, &de, inGrCol->handle(), inDataCol->handle(), current_stream()));
    return std::make_tuple(new column<t_out_gr_f>(g), new column<t_out_ext_f>(e));
}

/* nested_loop_join and left_semi_nto1_nested_loop_join (interfaces/join.h). */
template <class t_ve, class t_out_pos_l_f, class t_out_pos_r_f, class t_in_data_l_f, class t_in_data_r_f>
const std::tuple<const column<t_out_pos_l_f> *, const column<t_out_pos_r_f> *>
nested_loop_join(const column<t_in_data_l_f> *const inDataLCol, const column<t_in_data_r_f> *const inDataRCol,
                 const size_t outCountEstimate = 0)
{
    (void)outCountEstimate;
    const ms_format_desc dl = t_out_pos_l_f::desc(), dr = t_out_pos_r_f::desc();
    ms_col *l = nullptr, *r = nullptr;
    ms_check(ms_join(&l, &r, &dl, &dr, inDataLCol->handle(), inDataRCol->handle(), current_stream()));
    return std::make_tuple(new column<t_out_pos_l_f>(l), new column<t_out_pos_r_f>(r));
}

template <class t_ve, class t_out_pos_l_f, class t_in_data_l_f, class t_in_data_r_f>
const column<t_out_pos_l_f> *left_semi_nto1_nested_loop_join(const column<t_in_data_l_f> *const inDataLCol,
                                                             const column<t_in_data_r_f> *const inDataRCol,
                                                             const size_t outCountEstimate = 0)
{
    (void)outCountEstimate;
    const ms_format_desc d = t_out_pos_l_f::desc();
    ms_col *c = nullptr;
    ms_check(ms_semi_join(&c, &d, inDataLCol->handle(), inDataRCol->handle(), current_stream()));
    return new column<t_out_pos_l_f>(c);
}

template <class t_ve, class t_out_data_f, class t_in_data_f, class t_in_pos_f>
struct my_project_wit_t {
    static const column<t_out_data_f> *apply(const column<t_in_data_f> *const inDataCol,
                                             const column<t_in_pos_f> *const inPosCol)
    {
        const ms_format_desc d = t_out_data_f::desc();
        ms_col *c = nullptr;
        ms_check(ms_project(&c, &d, inDataCol->handle(), inPosCol->handle(), t_ve::lanes, current_stream()));
        return new column<t_out_data_f>(c);
    }
};

template <class t_ve, class t_in_f>
const column<uncompr_f> *agg_sum(const column<t_in_f> *const inDataCol)
{
    ms_col *c = nullptr;
    ms_check(ms_agg_sum(&c, inDataCol->handle(), current_stream()));
    return new column<uncompr_f>(c);
}

template <class t_ve, class t_out_pos_f, class t_in_pos_l_f, class t_in_pos_r_f>
const column<t_out_pos_f> *intersect_sorted(const column<t_in_pos_l_f> *const inPosLCol,
                                            const column<t_in_pos_r_f> *const inPosRCol,
                                            const size_t outPosCountEstimate = 0)
{
    (void)outPosCountEstimate;
    const ms_format_desc d = t_out_pos_f::desc();
    ms_col *c = nullptr;
    ms_check(ms_intersect_sorted(&c, &d, inPosLCol->handle(), inPosRCol->handle(), current_stream()));
    return new column<t_out_pos_f>(c);
}

/* intersect_compr.h:196-200: the compressed-left intersection; one kernel
 * path serves both. */
template <class t_ve, class t_out_pos_f, class t_in_pos_l_f, class t_in_pos_r_f>
struct my_intersect_wit_t {
    static const column<t_out_pos_f> *apply(const column<t_in_pos_l_f> *const inPosLCol,
                                            const column<t_in_pos_r_f> *const inPosRCol,
                                            const size_t outPosCountEstimate = 0)
    {
        return intersect_sorted<t_ve, t_out_pos_f>(inPosLCol, inPosRCol, outPosCountEstimate);
    }
};

template <class t_ve, class t_out_pos_f, class t_in_pos_l_f, class t_in_pos_r_f>
const column<t_out_pos_f> *merge_sorted(const column<t_in_pos_l_f> *const inPosLCol,
                                        const column<t_in_pos_r_f> *const inPosRCol)
{
    const ms_format_desc d = t_out_pos_f::desc();
    ms_col *c = nullptr;
    ms_check(ms_merge_sorted(&c, &d, inPosLCol->handle(), inPosRCol->handle(), current_stream()));
    return new column<t_out_pos_f>(c);
}

/* Binary operation codes for calc_binary's functor template (calc.h:102-119:
 * std::plus etc.). */
template <template <typename> class t_op> struct calc_op;
template <> struct calc_op<std::plus> { static constexpr uint32_t code = MS_OP_ADD; };
template <> struct calc_op<std::minus> { static constexpr uint32_t code = MS_OP_SUB; };
template <> struct calc_op<std::multiplies> { static constexpr uint32_t code = MS_OP_MUL; };
template <> struct calc_op<std::divides> { static constexpr uint32_t code = MS_OP_DIV; };
template <> struct calc_op<std::modulus> { static constexpr uint32_t code = MS_OP_MOD; };
template <> struct calc_op<std::bit_and> { static constexpr uint32_t code = MS_OP_AND; };
template <> struct calc_op<std::bit_or> { static constexpr uint32_t code = MS_OP_OR; };
template <> struct calc_op<std::bit_xor> { static constexpr uint32_t code = MS_OP_XOR; };

template <template <typename> class t_binary_op, class t_ve, class t_out_data_f, class t_in_data_l_f,
          class t_in_data_r_f>
const column<t_out_data_f> *calc_binary(const column<t_in_data_l_f> *const inDataLCol,
                                        const column<t_in_data_r_f> *const inDataRCol)
{
    const ms_format_desc d = t_out_data_f::desc();
    ms_col *c = nullptr;
    ms_check(ms_calc_binary(&c, calc_op<t_binary_op>::code, &d, inDataLCol->handle(), inDataRCol->handle(),
                            current_stream()));
    return new column<t_out_data_f>(c);
}

}  // namespace morphstore

#endif /* MORPHSTORE_HIP_HPP */
