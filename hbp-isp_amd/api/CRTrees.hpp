// Drop-in for CRTrees.hpp of HBP-ISP (CRTrees.hpp:14-78): labelling of the 1-NN functional
// graph on device arrays, backed by hbp_crtrees_get_clus / hbp_crtrees_compact.
#ifndef _CRTREES_HPP_
#define _CRTREES_HPP_

#include "hbp_api_common.hpp"

// CRTrees.cu:14-26: host read of the count after an exclusive scan
int get_num_from_scan(const int *predicate_d, const int *pos_scan_d, const int size);

class CRTrees {
protected:
    const int size;

public:
    explicit CRTrees(const int size) : size(size) {}
    ~CRTrees() = default;

    // CRTrees::get_clus (CRTrees.cu:315-354); the cycle mask (show_cycle) is not provided
    void get_clus(int *nn_d, int *clus_d, const int N, int *mask_cycle_d = nullptr);
    // CRTrees::compact_clus_label (CRTrees.cu:387-402)
    int compact_clus_label(int *clus_d, const int N);
};

#endif
