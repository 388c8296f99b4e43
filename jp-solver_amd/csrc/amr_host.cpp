// AMR orchestration: placeholder until the AMR kernels land.
#include "ctx.hpp"
namespace jp {
int amr_init(jp_ctx* c) { return fail(c, JP_ERR_STATE, "AMR not built"); }
int amr_base_step(jp_ctx* c, hipStream_t) { return fail(c, JP_ERR_STATE, "AMR not built"); }
int amr_composite(jp_ctx* c, void*) { return fail(c, JP_ERR_STATE, "AMR not built"); }
}  // namespace jp
