/*
 * selftest.h - The built-in test of the command line (the reference's src/cli/selftest.h:19).
 */
#pragma once

// Runs the GPU cases and the checker's planted-error cases; returns the number of failures.
int runSelfTest(bool verbose);
