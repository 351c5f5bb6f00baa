// Line-oriented driver of the host BigInt/BigFloat API for the test suite.
// Reads commands from stdin, one per line, and prints one result line each.  Integers are
// signed hex ("-1f", "0"); a BigFloat is printed as "<sign> <hex magnitude> <point_pos>".
//
//   mul A B | sq A | add A B | sub A B | muli A k | divi A k | shl A k | shr A k
//   and A B | or A B | xor A B | not A len | cmp A B | inc A | dec A
//   divmod A B mode          -> "Q R"  (mode 0 Truncate, 1 Floor, 2 Ceil, 3 RoundHalfUp)
//   todec A | fromdec S | fromhex S
//   fmul A pa B pb prec      -> BigFloat(A) >> 28*pa times BigFloat(B) >> 28*pb at prec limbs
//   fadd A pa B pb | fsub A pa B pb
//   frecip A pa prec | fprint A pa digits | ftodouble A pa | ffromdouble X (hex bits)
//   fround A pa mode prec rel
//   counts                   -> "<gpu products> <host products>"
//   poolinit N | pool        -> init_thread_pool(N) ("ok") | thread_pool_size()
//   decthr N                 -> sets DEC_STRING_BRUTE_THRESHOLDS[1] (limbs) and prints "ok"
//   devmuladd A B C D        -> A*B + C*D and A*A through DeviceInt (backend memory): "S SQ"
//   batch N SEED MAXBITS     -> checksum of BigInt::mul_batch over N SplitMix64 pairs (see below)
#include <bit>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "bigint/bigint.h"
#include "bigint/device.h"

using namespace bigint;

static BigInt H(const std::string& s) { return BigInt(s, true); }
static std::string X(const BigInt& x) { return x.to_string(true); }
static BigFloat F(const std::string& a, int64_t pa) { return BigFloat(H(a), -28 * pa); }
static std::string XF(const BigFloat& f) {
    // magnitude limbs -> BigInt through the public API: BigInt(f << 28*point_pos)
    BigFloat m = f;
    m.remove_sign();
    m <<= 28 * m.get_point_pos();
    BigInt mag(m, RoundMode::Truncate);
    std::ostringstream o;
    o << (f.sign() < 0 ? "-" : "+") << " " << X(mag) << " " << f.get_point_pos();
    return o.str();
}

int main() {
    std::ios::sync_with_stdio(false);
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string op;
        in >> op;
        std::string a, b;
        std::ostringstream out;
        try {
            if (op == "mul") {
                in >> a >> b;
                out << X(H(a) * H(b));
            } else if (op == "sq") {
                in >> a;
                BigInt x = H(a);
                x *= x;
                out << X(x);
            } else if (op == "add") {
                in >> a >> b;
                out << X(H(a) + H(b));
            } else if (op == "sub") {
                in >> a >> b;
                out << X(H(a) - H(b));
            } else if (op == "muli" || op == "divi" || op == "shl" || op == "shr") {
                long long k;
                in >> a >> k;
                BigInt x = H(a);
                if (op == "muli") x *= (int64_t)k;
                else if (op == "divi") x /= (int64_t)k;
                else if (op == "shl") x <<= (int64_t)k;
                else x >>= (int64_t)k;
                out << X(x);
            } else if (op == "and") {
                in >> a >> b;
                out << X(H(a) & H(b));
            } else if (op == "or") {
                in >> a >> b;
                out << X(H(a) | H(b));
            } else if (op == "xor") {
                in >> a >> b;
                out << X(H(a) ^ H(b));
            } else if (op == "not") {
                size_t len;
                in >> a >> len;
                BigInt x = H(a);
                x.bitwise_not(len);
                out << X(x);
            } else if (op == "cmp") {
                in >> a >> b;
                auto c = H(a) <=> H(b);
                out << (c < 0 ? -1 : c > 0 ? 1 : 0);
            } else if (op == "inc") {
                in >> a;
                BigInt x = H(a);
                ++x;
                out << X(x);
            } else if (op == "dec") {
                in >> a;
                BigInt x = H(a);
                --x;
                out << X(x);
            } else if (op == "divmod") {
                int mode;
                in >> a >> b >> mode;
                auto [q, r] = H(a).divmod(H(b), (RoundMode)mode);
                out << X(q) << " " << X(r);
            } else if (op == "todec") {
                in >> a;
                out << H(a).to_string(false);
            } else if (op == "fromdec") {
                in >> a;
                out << X(BigInt(a, false));
            } else if (op == "fromhex") {
                in >> a;
                out << X(BigInt(a, true));
            } else if (op == "fmul" || op == "fadd" || op == "fsub") {
                long long pa, pb;
                size_t prec = 0;
                in >> a >> pa >> b >> pb;
                if (op == "fmul") in >> prec;
                BigFloat fa = F(a, pa), fb = F(b, pb);
                if (op == "fmul") out << XF(BigFloat::mul(fa, fb, prec));
                else if (op == "fadd") out << XF(fa + fb);
                else out << XF(fa - fb);
            } else if (op == "fmuli" || op == "fshift") {  // BigFloat *= int64, BigFloat <<= bits
                long long pa, k;
                in >> a >> pa >> k;
                BigFloat fa = F(a, pa);
                if (op == "fmuli") fa *= (int64_t)k;
                else fa <<= (int64_t)k;
                out << XF(fa);
            } else if (op == "fsqr") {  // one object as both operands: the squaring route
                long long pa;
                size_t prec;
                in >> a >> pa >> prec;
                BigFloat fa = F(a, pa);
                out << XF(BigFloat::mul(fa, fa, prec));
            } else if (op == "frecip") {
                long long pa;
                size_t prec;
                in >> a >> pa >> prec;
                out << XF(F(a, pa).reciprocal(prec));
            } else if (op == "fprint") {
                long long pa;
                size_t digits;
                in >> a >> pa >> digits;
                out << F(a, pa).to_string(digits);
            } else if (op == "ftodouble") {
                long long pa;
                in >> a >> pa;
                double d = F(a, pa).to_double();
                out << std::hex << std::bit_cast<uint64_t>(d);
            } else if (op == "ffromdouble") {
                uint64_t bits;
                in >> std::hex >> bits;
                out << XF(BigFloat(std::bit_cast<double>(bits)));
            } else if (op == "fround") {
                long long pa, prec;
                int mode, rel;
                in >> a >> pa >> mode >> prec >> rel;
                BigFloat f = F(a, pa);
                f.round((RoundMode)mode, prec, (RoundRelativeTo)rel);
                out << XF(f);
            } else if (op == "batch") {
                // pair i: a, b of 1 + next % MAXBITS bits (top bit set) from SplitMix64(SEED);
                // every 7th pair squares a.  Checksum: sum (i+1) * (product mod 2^61-1).
                size_t N;
                uint64_t seed, maxbits;
                in >> N >> seed >> maxbits;
                uint64_t st = seed;
                auto next = [&st] {
                    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
                    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
                    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
                    return z ^ (z >> 31);
                };
                auto operand = [&] {
                    const uint64_t bits = 1 + next() % maxbits;
                    std::string hex;
                    for (uint64_t w = 0; w < (bits + 63) / 64; ++w) {
                        uint64_t v = next();
                        const uint64_t top = bits - 64 * w;
                        if (top <= 64) {
                            if (top < 64) v &= (1ull << top) - 1;
                            v |= 1ull << (top - 1);
                        }
                        char buf[17];
                        std::snprintf(buf, sizeof buf, "%016llx", (unsigned long long)v);
                        hex = std::string(buf) + hex;
                    }
                    return H(hex);
                };
                std::vector<BigInt> A, B;
                A.reserve(N);
                B.reserve(N);
                for (size_t i = 0; i < N; ++i) {
                    A.push_back(operand());
                    B.push_back(i % 7 == 0 ? BigInt(0) : operand());
                }
                std::vector<std::pair<const BigInt*, const BigInt*>> pairs(N);
                for (size_t i = 0; i < N; ++i) pairs[i] = {&A[i], i % 7 == 0 ? &A[i] : &B[i]};
                const std::vector<BigInt> P = BigInt::mul_batch(pairs, 1);
                const unsigned __int128 M = (((unsigned __int128)1) << 61) - 1;
                unsigned __int128 sum = 0;
                for (size_t i = 0; i < N; ++i) {
                    // product mod 2^61 - 1 from its hex digits
                    const std::string h = X(P[i]);
                    unsigned __int128 r = 0;
                    for (char c : h) r = (r * 16 + (uint64_t)(c <= '9' ? c - '0' : c - 'a' + 10)) % M;
                    sum = (sum + r * (i + 1)) % M;
                }
                out << (uint64_t)sum;
            } else if (op == "devmuladd") {
                std::string c, d;
                in >> a >> b >> c >> d;
                const BigInt A = H(a), B = H(b), C = H(c), D = H(d);
                auto dv = DeviceInt::upload({&A, &B, &C, &D});
                auto pr = DeviceInt::mul_batch({{&dv[0], &dv[1]}, {&dv[2], &dv[3]}, {&dv[0], &dv[0]}});
                auto sm = DeviceInt::add_batch({{&pr[0], &pr[1]}});
                auto back = DeviceInt::download({&sm[0], &pr[2]});
                out << X(back[0]) << " " << X(back[1]);
            } else if (op == "decthr") {
                size_type t;
                in >> t;
                BigInt::DEC_STRING_BRUTE_THRESHOLDS[1] = t;
                out << "ok";
            } else if (op == "poolinit") {  // init_thread_pool(N)
                uint32_t k;
                in >> k;
                init_thread_pool(k);
                out << "ok";
            } else if (op == "pool") {  // the pool's thread count (fixes it at the default if unset)
                out << thread_pool_size();
            } else if (op == "counts") {
                out << mul::gpu_product_count() << " " << mul::host_product_count();
            } else {
                out << "ERR unknown op";
            }
        } catch (const std::exception& e) {
            out << "EXC " << e.what();
        }
        std::cout << out.str() << "\n" << std::flush;
    }
    return 0;
}
