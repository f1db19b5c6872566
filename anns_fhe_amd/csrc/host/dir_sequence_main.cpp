// dir_sequence_main.cpp -- main_sim's start-up (cpp_host_model/src/main_sim.cpp:33-50)
// over the C-ABI: an FheDevice with the SimMemoryBus default arena sizes, the
// context, relinearisation key and rotation keys read from a test-vector directory
// by FHEHostDriver::initialize_context_from_directory, OP 99, the query inputs
// loaded from files, then eval_sub_plain -> eval_mult_relin_rescale ->
// eval_rotate(+1), and eval_rotate_single_step(-1) on the product
// (interactive_query_manager.cpp:85-119 slot order).
//
// usage: dir_sequence <tv_dir> <out_dir>
//   reads  <tv_dir>/ct_query_c0.bin, ct_query_c1.bin, pt_query.bin (L*N words each)
//   writes <out_dir>/{sub,mult,rot,rotm1}_c{0,1}.bin and rotation_keys.txt
//          (one line per loaded key: step auto_idx galois auto_map_offset evk_offset key_limbs)
#include <cstdio>
#include <string>

#include "fhe_host_driver.hpp"

using fhea_host::FHEHostDriver;
using fhea_host::FheDevice;

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s tv_dir out_dir\n", argv[0]);
        return 2;
    }
    const std::string tv = argv[1], out = argv[2];
    try {
        FheDevice mem_bus;
        FHEHostDriver driver(&mem_bus);
        driver.initialize_context_from_directory(tv);
        driver.load_context_to_hardware();
        const uint32_t N = driver.N, L = driver.get_config().sizeQl;
        const uint64_t W = (uint64_t)L * N, W1 = (uint64_t)(L - 1) * N;
        mem_bus.load_poly_file(0, tv + "/ct_query_c0.bin", W);
        mem_bus.load_poly_file(W, tv + "/ct_query_c1.bin", W);
        mem_bus.load_poly_file(2 * W, tv + "/pt_query.bin", W);
        driver.eval_sub_plain(0, W, 2 * W, 3 * W, 4 * W, L);
        driver.eval_mult_relin_rescale(3 * W, 4 * W, 3 * W, 4 * W, 5 * W, 6 * W, L);
        driver.eval_rotate(5 * W, 6 * W, 7 * W, 8 * W, 1, L - 1);
        driver.eval_rotate_single_step(5 * W, 6 * W, 9 * W, 10 * W, -1, L - 1);
        driver.sync();
        mem_bus.save_poly_file(3 * W, out + "/sub_c0.bin", W);
        mem_bus.save_poly_file(4 * W, out + "/sub_c1.bin", W);
        mem_bus.save_poly_file(5 * W, out + "/mult_c0.bin", W1);
        mem_bus.save_poly_file(6 * W, out + "/mult_c1.bin", W1);
        mem_bus.save_poly_file(7 * W, out + "/rot_c0.bin", W1);
        mem_bus.save_poly_file(8 * W, out + "/rot_c1.bin", W1);
        mem_bus.save_poly_file(9 * W, out + "/rotm1_c0.bin", W1);
        mem_bus.save_poly_file(10 * W, out + "/rotm1_c1.bin", W1);
        FILE *f = std::fopen((out + "/rotation_keys.txt").c_str(), "w");
        if (!f) return 3;
        for (const auto &kv : driver.rotation_keys()) {
            const auto &m = kv.second;
            std::fprintf(f, "%d %u %llu %u %llu %u\n", m.step, m.auto_idx, (unsigned long long)m.galois,
                         m.auto_map_offset, (unsigned long long)m.evk_offset, m.key_limbs);
        }
        std::fprintf(f, "relin %llu %u\n", (unsigned long long)driver.get_config().relin_key_offset,
                     driver.get_config().relin_key_limbs);
        std::fclose(f);
        std::printf("dir sequence done: N=%u L=%u\n", N, L);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "Simulation Error: %s\n", e.what());
        return 1;
    }
    return 0;
}
