/*
 * oc_main.c - fvc_cpu: run a flowVC deck through the oracle (CPU FP64 restatement).
 * TEST INFRASTRUCTURE ONLY.  Usage: fvc_cpu <deck.in> [nproc]
 */
#include "fvc_oracle.h"

#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <deck.in> [nproc]\n", argv[0]);
        return 2;
    }
    char msg[512] = {0};
    int rc = oc_run_deck(argv[1], argc > 2 ? atoi(argv[2]) : 0, msg, sizeof msg);
    if (rc) {
        fprintf(stderr, "fvc_cpu: %s (code %d)\n", msg, rc);
        return 1;
    }
    return 0;
}
