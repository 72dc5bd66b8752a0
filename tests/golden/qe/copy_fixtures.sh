#!/bin/sh
# Copies the pw.x outputs that the reference's test-suite keeps (data files,
# test-suite/testNNN/ref/pw.xml) into this folder as qe_testNNN.xml.
set -e
here=$(dirname "$0")
for t in test001 test004 test008 test014; do
  cp /root/reference/test-suite/$t/ref/pw.xml "$here/qe_$t.xml"
done
