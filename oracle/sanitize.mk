# Sanitizer builds of the CPU oracle (test infrastructure only): the checks of
# sanitize_check.c linked with sep_oracle.c, on the CPU.
#   _san/check_asan       AddressSanitizer + UndefinedBehaviorSanitizer, the parity variant
#   _san/check_asan_fast  the same over the contracted (timed CPU baseline) variant
#   _san/check_tsan       ThreadSanitizer (the threaded wave apply)
# make -f sanitize.mk   (from oracle/); run each binary, a report exits non-zero.
CC ?= gcc
SAN_COMMON = -O2 -g -fno-omit-frame-pointer -march=x86-64-v3 -fno-math-errno -std=c11 -Wall -Wno-unused-function
ASAN = -fsanitize=address,undefined -fno-sanitize-recover=all
TSAN = -fsanitize=thread
SRC = sanitize_check.c sep_oracle.c
DEPS = $(SRC) sanitize_check_impl.h sep_oracle_impl.h
LDLIBS = -lm -lpthread

all: _san/check_asan _san/check_asan_fast _san/check_tsan

_san:
	mkdir -p _san

_san/check_asan: $(DEPS) | _san
	$(CC) $(SAN_COMMON) $(ASAN) -ffp-contract=off -o $@ $(SRC) $(LDLIBS)

_san/check_asan_fast: $(DEPS) | _san
	$(CC) $(SAN_COMMON) $(ASAN) -ffp-contract=fast -DORC_PREFIX_FAST -o $@ $(SRC) $(LDLIBS)

_san/check_tsan: $(DEPS) | _san
	$(CC) $(SAN_COMMON) $(TSAN) -ffp-contract=off -o $@ $(SRC) $(LDLIBS)

.PHONY: all
