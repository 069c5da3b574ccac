/*
 * driver.h -- helpers shared by the C drivers of the sparse-matrix programs:
 * wall-clock seconds (as libgetput's wall_seconds, libgetput/libgetput.upc:433-438),
 * the version banner, and the models-mask bits (libgetput/libgetput.h:78-84).
 */
#ifndef BALE_DRIVER_H
#define BALE_DRIVER_H
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "bale_hip.h"

#define AGI_Model 1L
#define EXSTACK_Model 2L
#define EXSTACK2_Model 4L
#define CONVEYOR_Model 8L
#define ALTERNATE_Model 16L
#define ALL_Models 15L

static inline double wall_seconds(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* Initialise the GPU and print the banner and the command line. */
static inline void driver_start(int argc, char *argv[]) {
  if (bale_hip_init(0) != BALE_OK) {
    fprintf(stderr, "no GPU: %s\n", bale_hip_last_error());
    exit(1);
  }
  time_t now = time(NULL);
  struct tm *date = localtime(&now);
  fprintf(stderr, "Bale Version 2.1 (on one MI355X, libbale_hip): %04d-%02d-%02d.%02d:%02d\n",
          date->tm_year + 1900, date->tm_mon + 1, date->tm_mday, date->tm_hour, date->tm_min);
  fprintf(stderr, "Command:");
  for (int i = 0; i < argc; i++) fprintf(stderr, " %s", argv[i]);
  fprintf(stderr, "\n");
}

static inline void *must(void *p, const char *what) {
  if (!p) {
    fprintf(stderr, "ERROR: %s: %s\n", what, bale_hip_last_error());
    exit(1);
  }
  return p;
}
#endif
