/*
 * gviz_side.c -- the side table of gviz_side.h: open addressing over the
 * state addresses, linear probing, tombstones on removal, resized at half
 * load; one mutex (the reference's embedders are not thread-safe per state,
 * but distinct states may live on distinct threads).
 */
#include "gviz_side.h"

#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  const void *key; /* NULL: empty; TOMB: removed */
  void *value;
} slot;

static const char tomb_mark;
#define TOMB ((const void *)&tomb_mark)

static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
static slot *slots;
static size_t cap, used; /* used counts live keys and tombstones */

static size_t hash_of(const void *key) {
  uint64_t x = (uint64_t)(uintptr_t)key;
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  return (size_t)x;
}

static slot *find(const void *key) {
  if (!cap)
    return NULL;
  for (size_t i = hash_of(key) & (cap - 1), k = 0; k < cap; k++, i = (i + 1) & (cap - 1)) {
    if (slots[i].key == NULL)
      return NULL;
    if (slots[i].key == key)
      return &slots[i];
  }
  return NULL;
}

static int grow(void) {
  size_t ncap = cap ? cap * 2 : 64;
  slot *n = calloc(ncap, sizeof(slot));
  if (!n)
    return -1;
  size_t live = 0;
  for (size_t i = 0; i < cap; i++) {
    if (slots[i].key == NULL || slots[i].key == TOMB)
      continue;
    size_t j = hash_of(slots[i].key) & (ncap - 1);
    while (n[j].key)
      j = (j + 1) & (ncap - 1);
    n[j] = slots[i];
    live++;
  }
  free(slots);
  slots = n;
  cap = ncap;
  used = live;
  return 0;
}

void *gviz_side_get(const void *key) {
  pthread_mutex_lock(&mu);
  slot *s = key ? find(key) : NULL;
  void *v = s ? s->value : NULL;
  pthread_mutex_unlock(&mu);
  return v;
}

int gviz_side_put(const void *key, void *value, void **previous) {
  if (previous)
    *previous = NULL;
  if (!key)
    return -1;
  pthread_mutex_lock(&mu);
  slot *s = find(key);
  if (s) {
    if (previous)
      *previous = s->value;
    s->value = value;
    pthread_mutex_unlock(&mu);
    return 0;
  }
  if (2 * (used + 1) > cap && grow() != 0) {
    pthread_mutex_unlock(&mu);
    return -1;
  }
  size_t i = hash_of(key) & (cap - 1);
  while (slots[i].key && slots[i].key != TOMB)
    i = (i + 1) & (cap - 1);
  if (slots[i].key == NULL)
    used++;
  slots[i].key = key;
  slots[i].value = value;
  pthread_mutex_unlock(&mu);
  return 0;
}

void *gviz_side_take(const void *key) {
  pthread_mutex_lock(&mu);
  slot *s = key ? find(key) : NULL;
  void *v = NULL;
  if (s) {
    v = s->value;
    s->key = TOMB;
    s->value = NULL;
  }
  pthread_mutex_unlock(&mu);
  return v;
}
