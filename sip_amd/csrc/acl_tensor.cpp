// Minimal aclTensor descriptor (see include/acl_meta.h).
#include <stddef.h>

#include <new>
#include <vector>

#include "acl_meta.h"
#include "tensor_desc.h"

extern "C" {

aclTensor* aclCreateTensor(const int64_t* viewDims, uint64_t viewDimsNum, aclDataType dataType, const int64_t* stride,
                           int64_t offset, aclFormat format, const int64_t* storageDims, uint64_t storageDimsNum,
                           void* tensorData)
{
    if (viewDimsNum > 16 || (viewDimsNum > 0 && viewDims == nullptr) || offset < 0) return nullptr;
    if (storageDimsNum > 16 || (storageDimsNum > 0 && storageDims == nullptr)) return nullptr;
    aclTensor* t = new (std::nothrow) aclTensor();
    if (t == nullptr) return nullptr;
    try {
        t->view.assign(viewDims, viewDims + viewDimsNum);
        if (stride != nullptr) t->strides.assign(stride, stride + viewDimsNum);
        t->storage.assign(storageDims, storageDims + storageDimsNum);
    } catch (...) {
        delete t;
        return nullptr;
    }
    for (int64_t d : t->view) {
        if (d < 0) {
            delete t;
            return nullptr;
        }
    }
    t->dtype = dataType;
    t->format = format;
    t->offset = offset;
    t->data = tensorData;
    return t;
}

int32_t aclDestroyTensor(const aclTensor* tensor)
{
    delete tensor;
    return 0;
}

int32_t aclGetViewShape(const aclTensor* tensor, int64_t** viewDims, uint64_t* viewDimsNum)
{
    if (tensor == nullptr || viewDims == nullptr || viewDimsNum == nullptr) return 1;
    const size_t n = tensor->view.size();
    int64_t* dims = new (std::nothrow) int64_t[n > 0 ? n : 1];
    if (dims == nullptr) return 1;
    for (size_t i = 0; i < n; i++) dims[i] = tensor->view[i];
    *viewDims = dims;
    *viewDimsNum = n;
    return 0;
}

void aclFreeViewShape(int64_t* viewDims) { delete[] viewDims; }

int32_t aclGetDataType(const aclTensor* tensor, aclDataType* dataType)
{
    if (tensor == nullptr || dataType == nullptr) return 1;
    *dataType = tensor->dtype;
    return 0;
}

}  // extern "C"
