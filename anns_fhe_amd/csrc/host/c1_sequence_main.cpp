// c1_sequence_main.cpp -- the reference's main_sim.cpp flow in C++ over the
// C-ABI: eval_sub_plain -> eval_mult_relin_rescale -> eval_rotate(+1) at
// N=2^13, L=4, K=1, dnum=2 with seeded keys and uniform residues
// (cpp_host_model/src/main_sim.cpp:12-126, interactive_query_manager.cpp:85-119).
// Writes everything needed to check the run bit-exactly to one binary file:
//   u64 header[4] = {N, L, K, dnum}; primes[L+K]; inputs ct0|ct1|pt (3*L*N);
//   key words (relin then rot, 2 keys); sub (2*L*N); mult (2*(L-1)*N); rot (2*(L-1)*N)
#include <cstdio>
#include <random>
#include <vector>

#include "fhe_host_driver.hpp"

using fhea_host::FHEHostDriver;
using fhea_host::FheDevice;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
        return 2;
    }
    const uint32_t N = 8192, L = 4, K = 1, dnum = 2;
    const uint64_t W = (uint64_t)L * N, W1 = (uint64_t)(L - 1) * N;
    fhea_params p{};
    p.N = N;
    p.L = L;
    p.K = K;
    p.dnum = dnum;
    p.q_bits = 50;
    p.poly_words = 16 * W;
    p.key_words = 1 << 22;
    try {
        FheDevice dev(p);
        FHEHostDriver drv(&dev);
        std::vector<uint64_t> primes(L + K);
        fhea_ctx_primes(dev.ctx(), primes.data());
        std::vector<int8_t> s(N);
        fhea_host::throw_on(fhea_keygen_secret(dev.ctx(), 42, s.data()), "keygen_secret");
        const uint64_t ksz = fhea_key_size_words(dev.ctx(), L);
        fhea_host::throw_on(fhea_keygen_relin_key(dev.ctx(), s.data(), L, 43, 0), "relin key");
        const uint64_t g1 = fhea_galois_for_step(dev.ctx(), 1);
        fhea_host::throw_on(fhea_keygen_rotation_key(dev.ctx(), s.data(), g1, L, 44, ksz), "rot key");
        drv.set_relin_key(0, L);
        drv.add_rotation_key(1, g1, ksz, L);
        drv.load_context_to_hardware();

        std::mt19937_64 gen(42);
        std::vector<uint64_t> in(3 * W);
        for (int k = 0; k < 3; ++k)
            for (uint32_t l = 0; l < L; ++l)
                for (uint32_t j = 0; j < N; ++j) in[k * W + l * N + j] = gen() % primes[l];
        dev.write_poly(0, in.data(), in.size());
        // slots as in interactive_query_manager.cpp:88-105
        drv.eval_sub_plain(0, W, 2 * W, 3 * W, 4 * W, L);
        drv.eval_mult_relin_rescale(3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L);
        drv.eval_rotate(5 * W, 6 * W, 7 * W, 8 * W, 1, L - 1);
        drv.sync();

        std::vector<uint64_t> keys(2 * ksz), sub(2 * W), mult(2 * W1), rot(2 * W1);
        dev.read_key(0, keys.data(), keys.size());
        dev.read_poly(3 * W, sub.data(), 2 * W);
        dev.read_poly(5 * W, mult.data(), W1);
        dev.read_poly(6 * W, mult.data() + W1, W1);
        dev.read_poly(7 * W, rot.data(), W1);
        dev.read_poly(8 * W, rot.data() + W1, W1);
        FILE *f = std::fopen(argv[1], "wb");
        if (!f) return 3;
        uint64_t hdr[4] = {N, L, K, dnum};
        std::fwrite(hdr, 8, 4, f);
        std::fwrite(primes.data(), 8, primes.size(), f);
        std::fwrite(in.data(), 8, in.size(), f);
        std::fwrite(keys.data(), 8, keys.size(), f);
        std::fwrite(sub.data(), 8, sub.size(), f);
        std::fwrite(mult.data(), 8, mult.size(), f);
        std::fwrite(rot.data(), 8, rot.size(), f);
        std::fclose(f);
        std::printf("c1 sequence done: %s\n", argv[1]);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
