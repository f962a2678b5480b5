// Edge and face id assignment (first-occurrence order of the reference CPU
// builders) and the relations derived from the ids.
#pragma once

#include "relations.h"

namespace tg {

int ensure_vt_sorted(tg_mesh &m);
int ensure_edges(tg_mesh &m);  // VE, TE, EV (+ VV-aligned edge ids)
int ensure_faces(tg_mesh &m);  // VF, TF, FV
int ensure_et(tg_mesh &m);
int ensure_fe(tg_mesh &m);
int ensure_ft(tg_mesh &m);
int ensure_ef(tg_mesh &m);
int ensure_tt(tg_mesh &m);
int ensure_ff(tg_mesh &m);
int ensure_ee(tg_mesh &m);
int id_relation_view(tg_mesh &m, int32_t rel, RelView &r);

}  // namespace tg
