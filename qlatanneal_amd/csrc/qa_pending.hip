// Engines whose device path is not in this build yet: they report QA_EUNSUPPORTED.
#include "qa_internal.hpp"

namespace qa {
void sa_fp32(qa_ctx *, const qa_problem *, const qa_sa_params *, qa_sa_result *) {
    throw Error{QA_EUNSUPPORTED, "SA FP32 mode is not in this build; use the FP64 validation mode"};
}
void sqapt_validation(qa_ctx *, const qa_problem *, const qa_sqapt_params *, qa_sqapt_result *) {
    throw Error{QA_EUNSUPPORTED, "SQA-PT validation mode is not in this build"};
}
void sqapt_fp32(qa_ctx *, const qa_problem *, const qa_sqapt_params *, qa_sqapt_result *) {
    throw Error{QA_EUNSUPPORTED, "SQA-PT is not in this build"};
}
void chi_validation(qa_ctx *, const qa_problem *, const qa_chi_params *, qa_chi_result *) {
    throw Error{QA_EUNSUPPORTED, "sqa_chi validation mode is not in this build"};
}
void chi_fp32(qa_ctx *, const qa_problem *, const qa_chi_params *, qa_chi_result *) {
    throw Error{QA_EUNSUPPORTED, "sqa_chi is not in this build"};
}
}  // namespace qa
