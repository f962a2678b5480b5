// Placeholder: id-assigning relations are not built yet.
#include "ids.h"

namespace tg {
static int todo() {
    set_error("relation not implemented yet");
    return ERR_STATE;
}
int ensure_edges(tg_mesh &) { return todo(); }
int ensure_faces(tg_mesh &) { return todo(); }
int ensure_et(tg_mesh &) { return todo(); }
int ensure_fe(tg_mesh &) { return todo(); }
int ensure_ft(tg_mesh &) { return todo(); }
int ensure_ef(tg_mesh &) { return todo(); }
int ensure_tt(tg_mesh &) { return todo(); }
int ensure_ff(tg_mesh &) { return todo(); }
int ensure_ee(tg_mesh &) { return todo(); }
int id_relation_view(tg_mesh &, int32_t, RelView &) { return todo(); }
}  // namespace tg
