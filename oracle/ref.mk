# Compiles the reference's C constellation solver (preset 4) from
# /root/reference/10Bit_CUDA/004GCC_constellation.c, unchanged, into _ref/.
# ref_driver.c includes that file with its main() renamed and exposes its
# generation + counting for one N as a library call.  Nothing is copied.
REF ?= /root/reference/10Bit_CUDA/004GCC_constellation.c
CC ?= gcc

_ref/libref004.so: ref_driver.c $(REF)
	mkdir -p _ref
	$(CC) -O2 -fPIC -shared -w -DREF_SRC='"$(REF)"' -o $@ ref_driver.c
