// Internal interfaces between the HU kernels and the solver driver.
#pragma once
#include <vector>

#include "mtt_common.h"

namespace mtt {

// Owns device allocations of one solver handle; frees them all at the end.
class DeviceArena {
   public:
    DeviceArena() = default;
    DeviceArena(const DeviceArena&) = delete;
    DeviceArena& operator=(const DeviceArena&) = delete;
    ~DeviceArena() { release(); }
    void release() {
        for (void* p : ptrs_) (void)hipFree(p);
        ptrs_.clear();
        bytes_ = 0;
    }
    template <class T>
    int alloc(size_t count, T** out) {
        *out = nullptr;
        if (count == 0) count = 1;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
        ptrs_.push_back(p);
        bytes_ += count * sizeof(T);
        *out = static_cast<T*>(p);
        return MTT_OK;
    }
    template <class T>
    int upload(const std::vector<T>& v, T** out) {
        int rc = alloc<T>(v.size(), out);
        if (rc) return rc;
        if (!v.empty()) {
            hipError_t e = hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpy", __FILE__, __LINE__);
        }
        return MTT_OK;
    }
    size_t bytes() const { return bytes_; }

   private:
    std::vector<void*> ptrs_;
    size_t bytes_ = 0;
};

struct PlainTerminalPlan {
    int n_terms = 0, n_fold = 0, n_tiles = 0;
    int* d_terms = nullptr;          // terminal node ids
    int* d_fold_list = nullptr;      // indices into d_terms of fold terminals
    int* d_col_term = nullptr;       // showdown column -> terminal index
    uint8_t* d_col_player = nullptr; // showdown column -> player whose value it is
    void* d_tiles = nullptr;         // SdTile[n_tiles]
    int16_t* d_card_combos = nullptr;// [52][51] combos holding each card
};

int launch_hu_forward(const int* d_nodes_at_depth, int num, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                      const mtt_alive_combo* d_combos, const float* d_R, float* d_sigma, float* d_reach, int n_alive,
                      hipStream_t s, int action_stride = MTT_MAX_ACTIONS);
int launch_hu_backward(const int* d_nodes_at_depth, int num, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                       const mtt_alive_combo* d_combos, const float* d_sigma, const float* d_reach, float* d_cfv,
                       float* d_R, float* d_S, float w_pos, float w_neg, float w_strat, int n_alive, hipStream_t s,
                       int action_stride = MTT_MAX_ACTIONS);
int launch_hu_terminals_ref(const int* d_terms, int num, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                            const mtt_alive_combo* d_combos, const float* d_reach, const float* d_mats, float* d_cfv,
                            int n_alive, hipStream_t s);
int launch_hu_terminals_grouped(const PlainTerminalPlan& plan, const mtt_flat_node* d_nodes, const uint8_t* d_deck,
                                const mtt_alive_combo* d_combos, const float* d_reach, const float* d_mats,
                                long long* d_sums, float* d_cfv, int n_alive, hipStream_t s);
int build_terminal_plan(const std::vector<mtt_flat_node>& nodes, const std::vector<int>& terms, int n_alive,
                        const uint8_t* c1, const uint8_t* c2, PlainTerminalPlan& plan, DeviceArena& arena);

}  // namespace mtt
