"""Run-time format descriptors mirroring the reference's format templates.

The reference names a format by template arguments, e.g.
``dynamic_vbp_f<512, 2048, 8>`` (include/core/morphing/dynamic_vbp.h:83) or
``delta_f<1024, 8, dynamic_vbp_f<512, 64, 8>>`` (include/core/morphing/delta.h:78).
Here the same arguments fill an ``ms_format_desc`` (include/morphstore_hip.h).
"""
import ctypes

(MS_UNCOMPR, MS_STATIC_VBP, MS_DYNAMIC_VBP, MS_DELTA, MS_FOR, MS_RLE, MS_STATIC_VBP_PAD, MS_K_WISE_NS,
 MS_GROUP_SIMPLE, MS_BLOCKWISE_INDIVIDUAL) = range(10)
KIND_NAMES = {MS_UNCOMPR: "uncompr_f", MS_STATIC_VBP: "static_vbp_f", MS_DYNAMIC_VBP: "dynamic_vbp_f",
              MS_DELTA: "delta_f", MS_FOR: "for_f", MS_RLE: "rle_f", MS_STATIC_VBP_PAD: "static_vbp_f",
              MS_K_WISE_NS: "k_wise_ns_f", MS_GROUP_SIMPLE: "group_simple_f",
              MS_BLOCKWISE_INDIVIDUAL: "blockwise_individual_f"}

# format_code (include/core/format_selection/format_selector_commons.h:25-103);
# 1..64 are static_vbp_f at that bit width.  71 (rle_f pairs) is this
# library's own code.
(MS_BWI_UNCOMPR, MS_BWI_DYNAMIC_VBP, MS_BWI_GROUP_SIMPLE, MS_BWI_DELTA_DYNAMIC_VBP, MS_BWI_DELTA_GROUP_SIMPLE,
 MS_BWI_FOR_DYNAMIC_VBP, MS_BWI_FOR_GROUP_SIMPLE, MS_BWI_RLE) = (0, 65, 66, 67, 68, 69, 70, 71)
MS_BWI_SEL_BEST, MS_BWI_SEL_WORST, MS_BWI_SEL_STATIC = range(3)

MS_CMP_EQ, MS_CMP_NE, MS_CMP_LT, MS_CMP_LE, MS_CMP_GT, MS_CMP_GE = range(6)
(MS_OP_ADD, MS_OP_SUB, MS_OP_MUL, MS_OP_DIV, MS_OP_MOD, MS_OP_MIN, MS_OP_MAX,
 MS_OP_AND, MS_OP_OR, MS_OP_XOR) = range(10)

MS_OK, MS_ERR_INVALID, MS_ERR_NOMEM, MS_ERR_HIP, MS_ERR_SIZE, MS_ERR_CAPACITY, MS_ERR_NODEVICE, MS_ERR_RANGE = range(8)


class FormatDesc(ctypes.Structure):
    """ms_format_desc."""
    _fields_ = [(n, ctypes.c_uint32) for n in
                ("kind", "bw", "block", "page", "step",
                 "inner_kind", "inner_bw", "inner_block", "inner_page", "inner_step")]

    def name(self):
        k = self.kind
        if k == MS_UNCOMPR:
            return "uncompr_f"
        if k == MS_STATIC_VBP:
            return f"static_vbp_f<vbp_l<{self.bw},{self.step}>>"
        if k == MS_DYNAMIC_VBP:
            return f"dynamic_vbp_f<{self.block},{self.page},{self.step}>"
        if k == MS_RLE:
            return "rle_f"
        if k == MS_STATIC_VBP_PAD:
            return f"static_vbp_f<vbp_padding_l<{self.bw},{self.step}>>"
        if k == MS_K_WISE_NS:
            return f"k_wise_ns_f<{self.block}>"
        if k == MS_GROUP_SIMPLE:
            return f"group_simple_f<{self.step},uint64_t,{8 * self.step}>"
        if k == MS_BLOCKWISE_INDIVIDUAL:
            sel = ("exhaustive_size_format_selector_helper<true>", "exhaustive_size_format_selector_helper<false>",
                   f"static_format_selector_helper<{self.page}>")[self.bw] if self.bw <= 2 else f"selector{self.bw}"
            return f"blockwise_individual_f<{self.block},{sel}::selector>"
        inner = FormatDesc(self.inner_kind, self.inner_bw, self.inner_block, self.inner_page,
                           self.inner_step).name()
        if k == MS_DELTA:
            return f"delta_f<{self.block},{self.step},{inner}>"
        return f"for_f<{self.block},{self.page},{inner}>"

    __repr__ = name

    def block_size(self):
        """F::m_BlockSize (None for rle_f, which has none)."""
        if self.kind == MS_UNCOMPR:
            return 1
        if self.kind == MS_STATIC_VBP:
            return 64 * self.step
        if self.kind == MS_RLE:
            return None
        if self.kind == MS_STATIC_VBP_PAD:
            return 64 // self.bw * self.step
        if self.kind == MS_GROUP_SIMPLE:
            return self.step
        return self.block


class ColMeta(ctypes.Structure):
    """ms_col_meta (column_meta_data, include/core/storage/column_helper.h:32)."""
    _fields_ = [("count", ctypes.c_uint64), ("size_used", ctypes.c_uint64),
                ("size_compr", ctypes.c_uint64), ("size_alloc", ctypes.c_uint64)]


class BatchConfig(ctypes.Structure):
    """ms_batch_config."""
    _fields_ = [(n, ctypes.c_uint32) for n in ("wg_per_cu", "units_per_item", "tiles_per_item", "runs_per_item",
                                                "values_per_item", "waves_per_simd", "units_per_item_for",
                                                "units_per_item_narrow", "unroll")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def uncompr_f():
    return FormatDesc(MS_UNCOMPR)


def static_vbp_f(bw, step):
    """static_vbp_f<vbp_l<bw, step>>."""
    return FormatDesc(MS_STATIC_VBP, bw, 0, 0, step)


def static_vbp_padding_f(bw, step):
    """static_vbp_f<vbp_padding_l<bw, step>>."""
    return FormatDesc(MS_STATIC_VBP_PAD, bw, 0, 0, step)


def k_wise_ns_f(block=2):
    """k_wise_ns_f<block> (the reference implements block 2)."""
    return FormatDesc(MS_K_WISE_NS, 0, block, 0, 0)


def blockwise_individual_f(block, selector=MS_BWI_SEL_BEST, code=0):
    """blockwise_individual_f<block, selector, ve> with the 8-lane processing
    style's default formats; `code` is the fixed choice of MS_BWI_SEL_STATIC."""
    return FormatDesc(MS_BLOCKWISE_INDIVIDUAL, selector, block, code if selector == MS_BWI_SEL_STATIC else 0, 8)


def group_simple_f(lanes=8):
    """group_simple_f<lanes, uint64_t, 8 * lanes> (DEFAULT_GROUP_SIMPLE_F)."""
    return FormatDesc(MS_GROUP_SIMPLE, 0, 0, 0, lanes)


def dynamic_vbp_f(block, page, step):
    """dynamic_vbp_f<block, page, step>."""
    return FormatDesc(MS_DYNAMIC_VBP, 0, block, page, step)


def _with_inner(f, inner):
    f.inner_kind, f.inner_bw, f.inner_block = inner.kind, inner.bw, inner.block
    f.inner_page, f.inner_step = inner.page, inner.step
    return f


def delta_f(block, step, inner):
    """delta_f<block, step, inner>."""
    return _with_inner(FormatDesc(MS_DELTA, 0, block, 0, step), inner)


def for_f(block, page, inner):
    """for_f<block, page, inner>; the lane count is the inner format's step."""
    lanes = inner.step if inner.kind != MS_UNCOMPR else 1
    return _with_inner(FormatDesc(MS_FOR, 0, block, page, lanes), inner)


def rle_f():
    return FormatDesc(MS_RLE)


# DEFAULT_*_F(ve) for a vector extension of `lanes` 64-bit elements
# (include/core/morphing/default_formats.h:35-82).
def default_static_vbp_f(bw, lanes=8):
    return static_vbp_f(bw, lanes)


def default_dynamic_vbp_f(lanes=8):
    return dynamic_vbp_f(64 * lanes, 8 * lanes, lanes)


def default_delta_dynamic_vbp_f(lanes=8):
    return delta_f(1024, lanes, default_dynamic_vbp_f(lanes))


def default_for_dynamic_vbp_f(lanes=8):
    return for_f(1024, lanes, default_dynamic_vbp_f(lanes))
