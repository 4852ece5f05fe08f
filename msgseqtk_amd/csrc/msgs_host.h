/*
 * msgs_host.h -- host-side state of libmsgs_hip.so (context, resident database,
 * resident read batches) shared by its translation units.
 */
#ifndef MSGS_HOST_H
#define MSGS_HOST_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/msgs_hip.h"
#include "msgs_kernels.h"

struct DbState {
    int64_t n = 0;
    int64_t B[256] = {}, C[256] = {};
    double ll[16] = {};
    std::vector<std::string> names;
    std::vector<int64_t> chromLen, chromBD, tgtOff;
    int32_t nChrom = 0;
    /* device */
    uint8_t* dBd = nullptr;
    uint32_t* dSa = nullptr;
    uint32_t* dIsa = nullptr;
    uint8_t* dBwt = nullptr;
    OccBlock* dOcc = nullptr;
    int64_t* dNPos = nullptr;
    int64_t nN = 0;
    int64_t* dChromBD = nullptr;
    int64_t* dTgtOff = nullptr;
    int64_t* dChromLen = nullptr;
    uint8_t* dTgt = nullptr;
    std::vector<uint8_t> tgtHost;   /* chrom sequences + terminators (for record text) */
    void release();
    DevIndex view() const;
};

int buildIndexDevice(DbState& db, hipStream_t st, std::string& err);

#endif
