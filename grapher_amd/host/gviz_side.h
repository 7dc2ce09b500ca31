/*
 * gviz_side.h -- per-state extras of the gfx950 engine (device handles,
 * precision, residency flags, the CSR handed to the device), kept beside the
 * reference's state layouts in a table keyed by the state's address, so the
 * state structs keep the reference's fields and nothing else.
 */
#ifndef GVIZ_HOST_SIDE_H
#define GVIZ_HOST_SIDE_H

/* The extras of the state at `key`, or NULL. */
void *gviz_side_get(const void *key);
/* Attaches `value` to `key` (replacing and returning any earlier value).
 * Returns 0, or -1 when out of memory. */
int gviz_side_put(const void *key, void *value, void **previous);
/* Detaches and returns the extras of `key` (NULL if none). */
void *gviz_side_take(const void *key);

#endif
