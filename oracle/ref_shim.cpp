// ref_shim.cpp -- exposes a few of the reference's own hot-path routines,
// compiled from the reference headers where they lie, as plain C symbols so
// the tests can check the oracle against the real code.  TEST INFRASTRUCTURE
// ONLY: built into oracle/_ref/ by oracle/Makefile, never linked by the
// product.  Only headers that build with this image's toolchain alone are
// used (no spdlog / boost / CUDA):
//   alg/SpaceInnerProduct.hpp, alg/DiskIndex/RefineByComputation/RefineCPU.hpp,
//   alg/RankBoundRefinement/PruneCandidateByBound.hpp,
//   score_computation/CPUScoreTable.hpp, struct/VectorMatrix.hpp.
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "alg/SpaceInnerProduct.hpp"
#include "alg/DiskIndex/RefineByComputation/RefineCPU.hpp"
#include "alg/RankBoundRefinement/PruneCandidateByBound.hpp"
#include "score_computation/CPUScoreTable.hpp"
#include "struct/VectorMatrix.hpp"

extern "C" {

float ref_inner_product(const float *a, const float *b, int d) {
    return ReverseMIPS::InnerProduct(a, b, d);
}

int ref_refine_rank(const float *users, const float *items, int64_t n_user, int64_t n_item, int d,
                    float qip, int user_id) {
    ReverseMIPS::RefineCPU r(users, items, (uint64_t)n_user, (uint64_t)n_item, (uint64_t)d);
    int n_compute = 0;
    return r.RefineRank(qip, user_id, n_compute);
}

void ref_prune(const int *lb, const int *ub, int n_user, int topk, int *refine_user_size,
               int *n_result_user, int *n_prune_user, char *prune, char *result) {
    std::vector<int> lbv(lb, lb + n_user), ubv(ub, ub + n_user);
    std::vector<char> pv(prune, prune + n_user), rv(result, result + n_user);
    ReverseMIPS::PruneCandidateByBound(lbv, ubv, n_user, topk, *refine_user_size, *n_result_user,
                                       *n_prune_user, pv, rv);
    std::memcpy(prune, pv.data(), (size_t)n_user);
    std::memcpy(result, rv.data(), (size_t)n_user);
}

void ref_score_list(const float *users, const float *items, int64_t n_user, int64_t n_item, int d,
                    int start_user, int batch_n_user, float *out) {
    ReverseMIPS::CPUScoreTable t(users, items, (uint64_t)n_user, (uint64_t)n_item, (uint64_t)d);
    t.ComputeList(start_user, batch_n_user, out);
}

void ref_normalize(float *data, int n, int d) {
    std::unique_ptr<float[]> p = std::make_unique<float[]>((size_t)n * d);
    std::memcpy(p.get(), data, sizeof(float) * (size_t)n * d);
    ReverseMIPS::VectorMatrix m;
    m.init(p, n, d);
    m.vectorNormalize();
    std::memcpy(data, m.getRawData(), sizeof(float) * (size_t)n * d);
}

}  // extern "C"
