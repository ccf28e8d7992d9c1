// Host build of the shared Matrix Market text rules (csrc/mmtext.hpp) for the
// CPU tests: the value conversion and the line parser, by ctypes.
#include "../../spai-gmres-ir_amd/csrc/mmtext.hpp"

extern "C" {
// Returns the LineCode; *v the value, *fast 1 when no exact conversion was
// needed.
int mm_parse_value(const char *s, int n, int integer_field, double *v, int *fast) {
    double x = 0;
    const int c = bspai_mm::parse_value<false>(s, s + n, integer_field != 0, x);
    *fast = c != bspai_mm::kNeedsExact;
    if (c == bspai_mm::kNeedsExact) {
        const int c2 = bspai_mm::parse_value<true>(s, s + n, integer_field != 0, x);
        *v = x;
        return c2;
    }
    *v = x;
    return c;
}
int mm_parse_line(const char *s, int n, unsigned long long rows, unsigned long long cols,
                  int integer_field, unsigned long long *r, unsigned long long *c, double *v) {
    std::uint64_t rr = 0, cc = 0;
    double x = 0;
    const int code =
        bspai_mm::parse_entry_line<true>(s, s + n, rows, cols, integer_field != 0, rr, cc, x);
    *r = rr;
    *c = cc;
    *v = x;
    return code;
}
int mm_is_data_line(const char *s, int n) { return bspai_mm::is_data_line(s, s + n); }
}
