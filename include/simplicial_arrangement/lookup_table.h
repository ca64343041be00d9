// The lookup-table switches of the external simplicial_arrangement library,
// with the names its callers use (app/implicit_arrangement.cpp:29-42,
// app/material_interface.cpp:32-45, tests/test_implicit_networks.cpp:22,
// and the per-call switch at src/implicit_arrangement.cpp:176-199).
//
// Here the tables are built into the device library (one-function and
// two-function tables for arrangements, the two-material table for material
// interfaces), so load_lookup_table() has nothing to read from disk: it turns
// the tables on and returns true.  The switch is process-wide, as in the
// reference; it starts on.  The drop-in implicit_arrangement() and
// material_interface() use the tables when the switch is on and their
// use_lookup argument is true.  Results do not depend on it; speed does.
#pragma once

namespace simplicial_arrangement {

enum Mode { ARRANGEMENT = 0, MATERIAL_INTERFACE = 1 };

bool load_lookup_table(Mode mode = ARRANGEMENT);
void enable_lookup_table();
void disable_lookup_table();
// whether the switch is on (no counterpart in the reference)
bool lookup_table_enabled();

}  // namespace simplicial_arrangement
