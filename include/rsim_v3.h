/* The V3 wire format of librsim.so: RocketSim V3's arena dumps ("V3ARENA\0", version 2) and static dumps ("V3STATIC",
 * version 3) to and from the ArenaState and StaticData words that rsim.h steps. This is the data path a caller uses to
 * bring real V3 arenas onto the GPU and hand them back:
 *
 *   rsim_v3_parse_static(v3_static_dump)  -> StaticData  -> rsim_upload_world
 *   rsim_v3_parse_arena(v3_arena_dump)    -> ArenaState  -> rsim_upload_states -> rsim_sim8 ... -> rsim_download_states
 *   rsim_v3_write_arena(ArenaState, template)            -> a dump V3's importer takes back
 *
 * Replaces crates/rocketsim-cuda/src/dump.rs: parse_static :976-1011, parse_arena :885-900, write_arena :903-945,
 * arena_header :871-883, static_header :947-965, fnv1a64 :30-36, fold_records :395-421, unfold_records :425-447, and the
 * Template struct :205-238. Host code only (no GPU needed). A failed call returns RSIM_ERR_FORMAT (or RSIM_ERR_INVALID /
 * RSIM_ERR_SHORT) and rsim_last_error() holds the reason, in the reference's words.
 *
 * One difference from the reference, on purpose: a static dump's triangle, BVH node, leaf and static ray-cell entry
 * counts (header counts 3-6 and 11) may be anything up to StaticData's capacity (8,020 / 3,754 / 8,020 / 8,020 / 1,210),
 * where dump.rs:985-998 demands exactly those sizes; V3's own Soccar dumps fill them exactly and parse the same way. */
#ifndef RSIM_V3_H
#define RSIM_V3_H

#include "rsim.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RSIM_ERR_FORMAT (-5) /* a dump that does not parse, or a state the format cannot carry (rsim_last_error) */

#define RSIM_V3_ARENA_VERSION 2       /* dump.rs:16 */
#define RSIM_V3_STATIC_VERSION 3      /* dump.rs:17 */
#define RSIM_V3_ARENA_HEADER_BYTES 68  /* magic 8, version 4, config hash 8, 12 u32 counts (dump.rs:18) */
#define RSIM_V3_STATIC_HEADER_BYTES 76 /* magic 8, version 4, payload hash 8, payload length 8, 12 u32 counts (dump.rs:19) */
#define RSIM_V3_COUNTS 12
#define RSIM_V3_BODY_RECORD_BYTES 219  /* one BODY record of an arena dump (dump.rs:21) */
#define RSIM_V3_BODIES 23
#define RSIM_V3_RECORD_WORDS 12        /* one ContactRecord: is_swap, body a, body b, world point on b, on a, normal on b */

/* What an ArenaState leaves out of a V3 arena dump, taken from any dump of the same configuration by
 * rsim_v3_parse_arena and handed back to rsim_v3_write_arena (dump.rs:205-238): the raw records of the static bodies
 * (their companion_id bytes cleared; the dynamic bodies' entries are unused), the sweep-and-prune handle counters, the
 * manifold table stride, and the 24 bytes of Heatseeker / Dropshot ball info that Soccar never changes. */
typedef struct rsim_v3_template {
  uint64_t config_hash;
  uint8_t static_body[RSIM_V3_BODIES][RSIM_V3_BODY_RECORD_BYTES];
  uint8_t ball_other[24];
  uint16_t num_handles;
  uint16_t next_new_handle;
  uint32_t manifold_stride;
  uint32_t ball_other_set;
} rsim_v3_template;

uint64_t rsim_v3_template_bytes(void); /* sizeof(rsim_v3_template), for bindings that lay the struct out themselves */

/* FNV-1a, 64 bit (offset basis 0xcbf29ce484222325, prime 0x100000001b3): the static dump's payload hash. */
uint64_t rsim_fnv1a64(const void* bytes, uint64_t len);

/* The header of a static dump: magic, version 3, payload length and payload hash are checked. */
int rsim_v3_static_header(const void* bytes, uint64_t len, uint64_t* payload_hash, uint32_t counts[RSIM_V3_COUNTS]);
/* A static dump as StaticData (rsim_static_words() u32 words at static_out), table_sin filled as V3 fills it. */
int rsim_v3_parse_static(const void* bytes, uint64_t len, void* static_out);
/* StaticData as a static dump, through the same traversal the parser runs, sections in the order of V3's export_static
 * (no counterpart in dump.rs, which only reads static dumps). With cap below the size, writes nothing, sets *len_out to
 * the size and returns RSIM_ERR_SHORT. */
int rsim_v3_write_static(const void* static_data, void* out, uint64_t cap, uint64_t* len_out);

/* The header of an arena dump: magic and version 2 are checked. */
int rsim_v3_arena_header(const void* bytes, uint64_t len, uint64_t* config_hash, uint32_t counts[RSIM_V3_COUNTS]);
/* An arena dump (at a tick or a phase boundary) as an ArenaState (rsim_arena_words() words at state_out), in the world
 * static_data describes; tpl_out (may be null) receives what the state leaves out. */
int rsim_v3_parse_arena(const void* bytes, uint64_t len, const void* static_data, void* state_out, rsim_v3_template* tpl_out);
/* An ArenaState as an arena dump V3 imports. With cap below the size, writes nothing, sets *len_out to the size and
 * returns RSIM_ERR_SHORT. */
int rsim_v3_write_arena(const void* state, const void* static_data, const rsim_v3_template* tpl, void* out, uint64_t cap,
                        uint64_t* len_out);

/* V3's contact records of a tick (n records of RSIM_V3_RECORD_WORDS words, in call order) folded into the state's
 * fold_events, fold_car_normal and records, as the tick's P7 consumes them (dump.rs:395-421). */
int rsim_v3_fold_records(void* state, const void* static_data, const void* records, uint32_t n);
/* The inverse: records that fold to the same state (dump.rs:425-447). With cap below the count, writes nothing, sets
 * *n_out and returns RSIM_ERR_SHORT. */
int rsim_v3_unfold_records(const void* state, const void* static_data, void* records_out, uint32_t cap, uint32_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* RSIM_V3_H */
