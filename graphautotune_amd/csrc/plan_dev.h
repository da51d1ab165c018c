// Device program compiled from one or more Schedules. Plain data shared by
// the host compiler (plan.cpp) and the HIP engine (engine.hip).
//
// The program is a trie of nodes. A node at depth l chooses pattern vertex l
// from its loop slot (a prefix set), bounded above by the restriction
// vertices in `bound`, and then evaluates its ops (prefix sets that end at
// vertex l: ALIAS = a neighbour list, LIST = a materialised intersection,
// COUNT = the size of an intersection) and its terminals (the IEP
// combination that closes one pattern). Patterns whose first levels agree
// share the trie path, which is how batched plans share work.
#pragma once
#include <stdint.h>

#define GATP_MAX_DEPTH 8
#define GATP_MAX_NODES 128
#define GATP_MAX_SLOTS 96
#define GATP_MAX_OPS 160
#define GATP_MAX_TERMS 64
#define GATP_MAX_VIDS 256
#define GATP_MAX_SUBS 512
#define GATP_MAX_LEAF_OPS 8
#define GATP_MAX_STORE 12

enum { GATP_ALIAS = 0, GATP_LIST = 1, GATP_COUNT = 2 };
enum { GATP_VID_COUNT = 0, GATP_VID_LIST_REF = 1, GATP_VID_LIST_BOUNDED = 2 };

typedef struct GatOp {
    int16_t slot;      // output slot
    int16_t father;    // father slot (ALIAS or LIST) or -1: the op is N(v_l) itself
    int8_t kind;       // GATP_ALIAS / GATP_LIST / GATP_COUNT
    int8_t ngroups;    // truncation groups; 0 = no truncation
    uint8_t groups[4]; // bound = max over groups of min over levels in group
    uint8_t excl;      // COUNT: levels whose chosen vertex is excluded per element
    int8_t leaf_idx;   // COUNT evaluated per candidate at a leaf: counter index
    int8_t store;      // LIST: storage region
    int8_t pad;
} GatOp;

typedef struct GatNode {
    int8_t depth;
    uint8_t bound;     // candidates must be < min of v_f over these levels
    uint8_t excl;      // skip a candidate equal to v_j for these levels
    uint8_t ap_levels; // reference semantics: a candidate equal to v_j (j in
                       // mask) still descends, with this node's sets emptied
    int16_t parent;
    int16_t loop_slot; // -1 at the root
    int16_t op_begin, op_count;
    int16_t term_begin, term_count;
    int16_t child_begin, child_count; // children are consecutive node ids
    int16_t next_sibling;             // -1 when last child of its parent
    int8_t leaf;                      // no children
    int8_t pad;
} GatNode;

typedef struct GatVid {
    int16_t slot;
    int8_t mode;   // GATP_VID_*
    uint8_t bound; // LIST_BOUNDED: count elements < min v_f over these levels
    uint8_t excl;  // LIST_BOUNDED: distinct chosen vertices to subtract
    uint8_t refn;  // LIST_REF: subtract #{i < refn : v_i in list}, with multiplicity
    uint8_t pad[2];
} GatVid;

typedef struct GatSub {
    int16_t vid;   // index into vids (absolute)
    int16_t flag;  // closes a partition
    int32_t coef;
} GatSub;

typedef struct GatTerm {
    int16_t pattern;
    int16_t vid_begin, vid_count;
    int16_t sub_begin, sub_count;
    int16_t pad;
} GatTerm;

typedef struct GatSlot {
    int8_t kind;
    int8_t depth;     // level that computes it
    int8_t store;     // LIST storage region
    int8_t leaf_idx;  // COUNT at a leaf: per-candidate counter index, else -1
    int16_t node;
    int16_t pad;
} GatSlot;

typedef struct GatPlanDev {
    int32_t n_nodes, n_slots, n_ops, n_terms, n_vids, n_subs;
    int32_t n_patterns, n_store, max_leaf_ops, max_depth;
    int32_t root_slot;  // ALIAS slot of prefix [0]
    int32_t pad;
    GatNode nodes[GATP_MAX_NODES];
    GatSlot slots[GATP_MAX_SLOTS];
    GatOp ops[GATP_MAX_OPS];
    GatTerm terms[GATP_MAX_TERMS];
    GatVid vids[GATP_MAX_VIDS];
    GatSub subs[GATP_MAX_SUBS];
} GatPlanDev;
