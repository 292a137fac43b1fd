// kw/kwAsset.h -- the option record (reference: src/Core/kwAsset.h:12-45).
// Byte-identical to kwgpu_option, so a std::vector<kw::Option> goes to the GPU as is.
#pragma once

#include <cstddef>
#include <cstdio>
#include <utility>
#include <ostream>
#include <string>
#include <tuple>

#include "kw/kwTypes.h"
#include "kwgpu.h"

namespace kw {

// Field order, widths and names are the reference's (its callers and the 56-byte layout
// depend on them). Units: t in years, k and s in currency, z, r and q as annual
// fractions (0.2 is 20 %); e is 1 for early exercise, w is -1 for a put and +1 for a call.
struct Option {
    f64 t;
    f64 k;
    f64 z;
    f64 r;
    f64 q;
    f64 s;
    u8 e;
    i8 w;

    // ordering used to sort portfolios: (t, k, z, q, r, s, e, w) lexicographically
    friend bool operator<(const Option& l, const Option& r) {
        return std::tie(l.t, l.k, l.z, l.q, l.r, l.s, l.e, l.w) <
               std::tie(r.t, r.k, r.z, r.q, r.r, r.s, r.e, r.w);
    }

    // One-line description in the reference's format: exercise, then k q r s t with
    // six decimals (std::to_string's "%f"), then the parity, then z.
    std::string asString() const {
        char buf[256];
        std::snprintf(buf, sizeof(buf), "<Option e=%s k=%f q=%f r=%f s=%f t=%f w=%s z=%f>",
                      e ? "amer" : "euro", k, q, r, s, t, w > 0 ? "Call" : "Put", z);
        return buf;
    }
};

static_assert(sizeof(Option) == sizeof(kwgpu_option), "kw::Option must match kwgpu_option");
static_assert(offsetof(Option, e) == offsetof(kwgpu_option, e) &&
                  offsetof(Option, w) == offsetof(kwgpu_option, w) &&
                  offsetof(Option, s) == offsetof(kwgpu_option, s),
              "kw::Option must match kwgpu_option");

// Stream form (the reference's format: spot and expiry first, comma-separated, stream
// precision for the numbers).
inline std::ostream& operator<<(std::ostream& os, const Option& o) {
    os << "<Option s=" << o.s << " t=" << o.t;
    const std::pair<const char*, f64> rest[] = {{"k", o.k}, {"z", o.z}, {"r", o.r}, {"q", o.q}};
    for (const auto& [name, v] : rest) os << ", " << name << "=" << v;
    return os << ", " << (o.e ? "amer" : "euro") << ", " << (o.w > 0 ? "call" : "put") << ">";
}

}  // namespace kw
