# Sanitizer builds of the threaded CPU oracle (host code only, g++; test
# infrastructure, never shipped):
#
#   make -C oracle -f sanitize.mk      -> _san/tsan_check, _san/asan_check
#
# tsan_check: ThreadSanitizer; asan_check: AddressSanitizer + UBSan.
CXX      ?= g++
SANFLAGS  = -O1 -g -fno-omit-frame-pointer -ffp-contract=off -std=c++14 -pthread

all: _san/tsan_check _san/asan_check

_san/tsan_check: sanitize_check.cc dream_oracle.cc
	mkdir -p _san
	$(CXX) $(SANFLAGS) -fsanitize=thread -o $@ sanitize_check.cc dream_oracle.cc

_san/asan_check: sanitize_check.cc dream_oracle.cc
	mkdir -p _san
	$(CXX) $(SANFLAGS) -fsanitize=address,undefined -fno-sanitize-recover=undefined \
	    -o $@ sanitize_check.cc dream_oracle.cc

.PHONY: all
