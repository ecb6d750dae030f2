# Builds openPSTD's own copy of unqlite (1.1.6, unqlite/unqlite.c of the
# reference checkout) into oracle/_ref/, the checker of the .pstd database
# format: the tests write files with it and read them with the product, and
# the other way round.  Compiled from the sources where they lie; nothing is
# copied.  Runs only where the reference checkout exists.
REF ?= /root/reference
OUT := _ref

all: $(OUT)/libunqlite_ref.so

$(OUT)/libunqlite_ref.so: $(REF)/unqlite/unqlite.c $(REF)/unqlite/unqlite.h
	mkdir -p $(OUT)
	$(CC) -O1 -w -shared -fPIC -o $@.tmp $(REF)/unqlite/unqlite.c
	mv $@.tmp $@

.PHONY: all
