# Host-only AddressSanitizer + UndefinedBehaviorSanitizer build of the oracle
# with its driver (oracle/sanitize_main.c): make -f oracle/sanitize.mk
CC       = gcc
SANFLAGS = -O1 -g -std=c11 -Wall -Wextra -Wno-unused-parameter -fopenmp -fno-omit-frame-pointer \
           -fsanitize=address,undefined -fno-sanitize-recover=undefined
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))

all: $(HERE)_sanitize/fss_sanitize

$(HERE)_sanitize/fss_sanitize: $(HERE)aes.c $(HERE)fss.c $(HERE)sanitize_main.c $(HERE)fss_oracle.h
	mkdir -p $(HERE)_sanitize
	$(CC) $(SANFLAGS) -o $@ $(HERE)aes.c $(HERE)fss.c $(HERE)sanitize_main.c

clean:
	rm -rf $(HERE)_sanitize

.PHONY: all clean
