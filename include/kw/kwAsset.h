// kw/kwAsset.h -- the option record (reference: src/Core/kwAsset.h:12-45).
// Byte-identical to kwgpu_option, so a std::vector<kw::Option> goes to the GPU as is.
#pragma once

#include <cstddef>
#include <ostream>
#include <string>
#include <tuple>

#include "kw/kwTypes.h"
#include "kwgpu.h"

namespace kw {

struct Option {
    f64 t;  /// time to maturity     1.0 = 365 days
    f64 k;  /// strike               100 = $100
    f64 z;  /// volatility           0.2 = 20% per year
    f64 r;  /// interest rate        0.05 = 5% per year
    f64 q;  /// dividend rate        0.03 = 3% per year
    f64 s;  /// spot price           100 = $100
    u8 e;   /// early exercise       0 = European, 1 = American
    i8 w;   /// parity               -1 = put, +1 call

    friend bool operator<(const Option& l, const Option& r) {
        return std::tie(l.t, l.k, l.z, l.q, l.r, l.s, l.e, l.w) <
               std::tie(r.t, r.k, r.z, r.q, r.r, r.s, r.e, r.w);
    }

    std::string asString() const {
        std::string res = "<Option";
        res += e ? " e=amer" : " e=euro";
        res += " k=" + std::to_string(k);
        res += " q=" + std::to_string(q);
        res += " r=" + std::to_string(r);
        res += " s=" + std::to_string(s);
        res += " t=" + std::to_string(t);
        res += w > 0 ? " w=Call" : " w=Put";
        res += " z=" + std::to_string(z);
        res += ">";
        return res;
    }
};

static_assert(sizeof(Option) == sizeof(kwgpu_option), "kw::Option must match kwgpu_option");
static_assert(offsetof(Option, e) == offsetof(kwgpu_option, e) &&
                  offsetof(Option, w) == offsetof(kwgpu_option, w) &&
                  offsetof(Option, s) == offsetof(kwgpu_option, s),
              "kw::Option must match kwgpu_option");

inline std::ostream& operator<<(std::ostream& os, const Option& o) {
    return os << "<Option s=" << o.s << " t=" << o.t << ", k=" << o.k << ", z=" << o.z
              << ", r=" << o.r << ", q=" << o.q << ", " << (o.e ? "amer" : "euro") << ", "
              << (o.w > 0 ? "call" : "put") << ">";
}

}  // namespace kw
