#!/bin/sh
# Writes the device headers named on the command line as C++ initialisers
# {"basename", R"MINDRTC(<text>)MINDRTC"}, for the runtime compiler (rtc.hip).
for f in "$@"; do
  printf '{"%s", R"MINDRTC(' "$(basename "$f")"
  cat "$f"
  printf ')MINDRTC"},\n'
done
