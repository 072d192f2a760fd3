// Compiles and links a C++ caller of the C-ABI and the SpmvSpmmProblem
// mirror (include/spmvspmm.hpp); runs only calls that need no GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "spmvspmm.hpp"

int main() {
    // Host-side argument checks return before touching the device.
    SpmvSpmmProblem prob(4, 4, 0, 8, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    prob.run(DENSE_ROW_MAJOR, 3, ALG_CSR_SCALAR);               // prints, returns
    if (prob.benchmark(DENSE_ROW_MAJOR, 6, ALG_CSR_SCALAR) != -1.0f) return 1;
    if (dspmm_spmm(ALG_COO_VECTOR, DENSE_COL_MAJOR, 4, 4, 0, 5, nullptr, nullptr, nullptr,
                   nullptr, nullptr, nullptr, nullptr) != DSPMM_STATUS_NOT_SUPPORTED)
        return 2;
    if (std::strcmp(dspmm_get_error_string(DSPMM_STATUS_SUCCESS), "success") != 0) return 3;
    // CUArray::init_random: the reference's generator, host side only
    dspmm::CUArray<float> f;
    f.init_random(1000, 7);
    dspmm::CUArray<int> n;
    n.init_random(1000, 7);
    for (size_t i = 0; i < 1000; i++) {
        const float v = f.get_host_array()[i];
        const int k = n.get_host_array()[i];
        if (v < -0.5f || v > 0.499f || k < -50 || k > 49) return 4;
    }
    std::srand(7);
    const float first = (float)(std::rand() % 1000 - 500) / 1000;
    if (f.get_host_array()[0] != first) return 5;
    std::printf("ok %s\n", dspmm_version());
    // sizes of the public structs, compared with the Python mirror's
    std::printf("sizes device_info=%zu matrix_desc=%zu config=%zu batch_item=%zu batch_info=%zu "
                "server_config=%zu request=%zu completion=%zu coo=%zu server_stats=%zu\n",
                sizeof(dspmm_device_info), sizeof(dspmm_matrix_desc), sizeof(dspmm_config),
                sizeof(dspmm_batch_item), sizeof(dspmm_batch_info), sizeof(dspmm_server_config),
                sizeof(dspmm_request), sizeof(dspmm_completion), sizeof(dspmm_coo),
                sizeof(dspmm_server_stats));
    return 0;
}
