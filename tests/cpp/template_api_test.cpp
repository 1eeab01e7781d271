// The reference's query shape through the header-only template API
// (select_sum_benchmark.cpp:78-118 and the operator known-answer tests of
// test/core/operators/*_test.cpp), checked against brute force.  Prints
// "template api ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "morphstore_hip.hpp"

using namespace morphstore;
using ve = vectorlib::hip_gfx950;

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                                    \
        }                                                                \
    } while (0)

static uint64_t splitmix(uint64_t i)
{
    uint64_t z = 0x4D6F7270ull + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main()
{
    ms_check(ms_init(0));
    // select(x < 2^11) -> positions -> project -> agg_sum over static_vbp<12, 1>
    const size_t n = (1u << 20) + 77;
    std::vector<uint64_t> x(n);
    for (size_t i = 0; i < n; i++) x[i] = splitmix(i) >> 52;
    using sf = static_vbp_f<vbp_l<12, 1>>;
    const column<uncompr_f> *base = make_column(x);
    const column<sf> *c = morph<ve, sf>(base);
    CHECK(c->get_count_values() == n);
    CHECK(c->get_size_compr_byte() == (n / 64) * 64 * 12 / 8);
    CHECK(values_of(c) == x);
    const column<uncompr_f> *pos = my_select_wit_t<less, ve, uncompr_f, sf>::apply(c, 1u << 11);
    const column<uncompr_f> *dat = my_project_wit_t<ve, uncompr_f, sf, uncompr_f>::apply(c, pos);
    const column<uncompr_f> *sum = agg_sum<ve, uncompr_f>(dat);
    uint64_t want = 0, cnt = 0;
    std::vector<uint64_t> want_pos;
    for (size_t i = 0; i < n; i++)
        if (x[i] < (1u << 11)) { want += x[i]; cnt++; want_pos.push_back(i); }
    CHECK(values_of(pos) == want_pos);
    CHECK(values_of(sum) == std::vector<uint64_t>{want});
    // a same-format morph is the input itself (morph.h:128-151)
    CHECK(((const void *)morph<ve, sf>(c) == (const void *)c));
    // compressed positions through the write-iterator, cascades
    using df = delta_f<1024, 8, dynamic_vbp_f<512, 64, 8>>;
    using ff = for_f<1024, 8, dynamic_vbp_f<512, 64, 8>>;
    const column<df> *pd = my_select_wit_t<greaterequal, ve, df, sf>::apply(c, 100);
    std::vector<uint64_t> ge;
    for (size_t i = 0; i < n; i++)
        if (x[i] >= 100) ge.push_back(i);
    CHECK(values_of(pd) == ge);
    const column<ff> *cf = morph<ve, ff>(base);
    CHECK(values_of(cf) == x);
    // known answers: intersect / merge / calc (intersect_test.cpp, merge_test.cpp, calc_binary_test.cpp shapes)
    const column<uncompr_f> *l = make_column(std::vector<uint64_t>{1, 3, 5, 7, 9, 11});
    const column<uncompr_f> *r = make_column(std::vector<uint64_t>{3, 4, 5, 10, 11});
    const column<uncompr_f> *is = intersect_sorted<ve, uncompr_f>(l, r);
    CHECK(values_of(is) == (std::vector<uint64_t>{3, 5, 11}));
    const column<uncompr_f> *ms = merge_sorted<ve, uncompr_f>(l, r);
    CHECK(values_of(ms) == (std::vector<uint64_t>{1, 3, 4, 5, 7, 9, 10, 11}));
    const column<uncompr_f> *a = make_column(std::vector<uint64_t>{10, 20, 30});
    const column<uncompr_f> *b = make_column(std::vector<uint64_t>{1, 2, 3});
    const column<uncompr_f> *s = calc_binary<std::minus, ve, uncompr_f>(a, b);
    CHECK(values_of(s) == (std::vector<uint64_t>{9, 18, 27}));
    // rle round trip
    const column<rle_f> *rl = morph<ve, rle_f>(make_column(std::vector<uint64_t>{4, 4, 4, 9, 9, 1}));
    CHECK(rl->get_size_used_byte() == 48);
    CHECK(values_of(rl) == (std::vector<uint64_t>{4, 4, 4, 9, 9, 1}));
    // between, group, joins: the known answers of between/group/join tests
    const column<uncompr_f> *bt =
        my_between_wit_t<greaterequal, lessequal, ve, uncompr_f, sf>::apply(c, 100, 200);
    std::vector<uint64_t> wbt;
    for (size_t i = 0; i < n; i++)
        if (x[i] >= 100 && x[i] <= 200) wbt.push_back(i);
    CHECK(values_of(bt) == wbt);
    auto gu = group<ve, uncompr_f, uncompr_f>(make_column(std::vector<uint64_t>{333, 333, 111, 333, 222, 111}));
    CHECK(values_of(std::get<0>(gu)) == (std::vector<uint64_t>{0, 0, 1, 0, 2, 1}));
    CHECK(values_of(std::get<1>(gu)) == (std::vector<uint64_t>{0, 2, 4}));
    auto gb = group<ve, uncompr_f, uncompr_f>(std::get<0>(gu),
                                              make_column(std::vector<uint64_t>{77, 88, 77, 77, 88, 88}));
    CHECK(values_of(std::get<0>(gb)) == (std::vector<uint64_t>{0, 1, 2, 0, 3, 4}));
    auto jn = nested_loop_join<ve, uncompr_f, uncompr_f>(make_column(std::vector<uint64_t>{22, 44, 11, 22, 55, 77}),
                                                         make_column(std::vector<uint64_t>{33, 22, 22, 11}));
    CHECK(values_of(std::get<0>(jn)) == (std::vector<uint64_t>{0, 0, 2, 3, 3}));
    CHECK(values_of(std::get<1>(jn)) == (std::vector<uint64_t>{1, 2, 3, 1, 2}));
    const column<uncompr_f> *sj = left_semi_nto1_nested_loop_join<ve, uncompr_f>(
        make_column(std::vector<uint64_t>{11, 22, 33, 11, 44, 55}), make_column(std::vector<uint64_t>{22, 22, 33, 44, 33}));
    CHECK(values_of(sj) == (std::vector<uint64_t>{1, 2, 4}));
    // the other candidate formats and binary_io
    using pf = static_vbp_f<vbp_padding_l<12, 8>>;
    using gf = group_simple_f<8, uint64_t, 64>;
    using kf = k_wise_ns_f<2>;
    CHECK(values_of(morph<ve, pf>(base)) == x);
    CHECK(values_of(morph<ve, gf>(base)) == x);
    CHECK(values_of(morph<ve, kf>(base)) == x);
    const column<sf> *c2 = morph<ve, sf>(base);
    CHECK(c2->prepare_for_random_access<ve>() == (n % 64 != 0));
    binary_io<sf>::store(c2, "/tmp/ms_template_api_col.bin");
    const column<sf> *c3 = binary_io<sf>::load("/tmp/ms_template_api_col.bin");
    CHECK(values_of(c3) == x);
    // errors surface as exceptions
    bool threw = false;
    try {
        calc_binary<std::plus, ve, uncompr_f>(a, l);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    // blockwise_individual_f: the exhaustive best-size selector and a static
    // one, read back by decompression and by agg_sum
    {
        using bf = blockwise_individual_f<2048, exhaustive_size_format_selector_helper<true>::selector, ve>;
        using rf = blockwise_individual_f<2048, static_format_selector_helper<format_code::defaultForDynamicVbp>::selector,
                                          ve>;
        const column<bf> *cb = morph<ve, bf>(base);
        const column<rf> *cr = morph<ve, rf>(base);
        CHECK(values_of(cb) == x);
        CHECK(values_of(cr) == x);
        uint64_t total = 0;
        for (uint64_t v : x) total += v;
        CHECK(values_of(agg_sum<ve, bf>(cb)) == std::vector<uint64_t>{total});
        CHECK(values_of(agg_sum<ve, rf>(cr)) == std::vector<uint64_t>{total});
        CHECK(cr->get_size_compr_byte() >= cb->get_size_compr_byte());
    }
    delete pos; delete dat; delete sum; delete pd; delete cf; delete is; delete ms; delete s;
    delete rl; delete a; delete b; delete l; delete r; delete c; delete base;
    std::printf("template api ok: %zu values, %llu selected\n", n, (unsigned long long)cnt);
    return 0;
}
