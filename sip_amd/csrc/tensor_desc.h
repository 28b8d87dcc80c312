// Definition of the tensor descriptor behind the opaque aclTensor handle.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "acl_meta.h"

struct aclTensor {
    std::vector<int64_t> view;
    std::vector<int64_t> strides;  // empty: dense row-major
    std::vector<int64_t> storage;
    aclDataType dtype = ACL_DT_UNDEFINED;
    aclFormat format = ACL_FORMAT_ND;
    int64_t offset = 0;  // elements
    void* data = nullptr;
};

namespace asdfft {

inline int64_t element_bytes(aclDataType t)
{
    switch (t) {
        case ACL_FLOAT:
            return 4;
        case ACL_COMPLEX64:
            return 8;
        case ACL_DOUBLE:
            return 8;
        case ACL_COMPLEX128:
            return 16;
        case ACL_FLOAT16:
            return 2;
        default:
            return 0;
    }
}

// True when the view is dense row-major (strides absent or equal to the packed strides).
inline bool is_dense(const aclTensor* t)
{
    if (t->strides.empty()) return true;
    int64_t expect = 1;
    for (size_t i = t->view.size(); i-- > 0;) {
        if (t->view[i] != 1 && t->strides[i] != expect) return false;
        expect *= t->view[i];
    }
    return true;
}

inline void* data_address(const aclTensor* t)
{
    if (t->data == nullptr) return nullptr;
    return static_cast<char*>(t->data) + t->offset * element_bytes(t->dtype);
}

inline int64_t numel(const aclTensor* t)
{
    int64_t n = 1;
    for (int64_t d : t->view) n *= d;
    return n;
}

}  // namespace asdfft
