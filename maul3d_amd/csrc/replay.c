// The M3J1 replay container: a snapshot plus a journal plus the recorder's
// final hash, framed in memory (include/maul3d/replay.h:15-62,
// src/replay.c).  The framing is the reference's, so a session recorded on
// the MI355X build is verified by the reference's tools and the other way
// round:
//
//   magic "M3J1" u32 | version u32 (1) | snapshotBytes i32 | journalBytes i32 |
//   stepCount i32 | opCount i32 | finalHash u64 | snapshot | journal
//
// Nothing here touches a world: the embedded snapshot validates itself when
// it is restored, and the container guards only its own framing.
#include "maul3d_amd.h"

#include <string.h>

#define M3_REPLAY_MAGIC 0x314A334Du
#define M3_REPLAY_VERSION 1u
#define M3_REPLAY_HEADER 32
#define M3_OP_STEP 1

bool m3JournalDescribe(const void* journal, int32_t bytes, m3JournalInfo* out)
{
    if (journal == NULL || bytes < 0 || out == NULL)
    {
        return false;
    }
    const uint8_t* p = (const uint8_t*)journal;
    int32_t ops = 0;
    int32_t steps = 0;
    for (int32_t at = 0; at < bytes;)
    {
        int32_t record[2];
        if (bytes - at < 8)
        {
            return false; // a record header cut short
        }
        memcpy(record, p + at, 8);
        if (record[0] <= 0 || record[1] < 0 || record[1] > bytes - at - 8)
        {
            return false;
        }
        ops += 1;
        steps += record[0] == M3_OP_STEP;
        at += 8 + record[1];
    }
    out->opCount = ops;
    out->stepCount = steps;
    return true;
}

int32_t m3ReplayEncodeSize(int32_t snapshotBytes, int32_t journalBytes)
{
    if (snapshotBytes <= 0 || journalBytes < 0 ||
        (int64_t)M3_REPLAY_HEADER + snapshotBytes + journalBytes > 0x7FFFFFFF)
    {
        return 0;
    }
    return M3_REPLAY_HEADER + snapshotBytes + journalBytes;
}

int32_t m3ReplayEncode(const void* snapshot, int32_t snapshotBytes, const void* journal,
                       int32_t journalBytes, uint64_t finalHash, void* out, int32_t capacity)
{
    const int32_t need = m3ReplayEncodeSize(snapshotBytes, journalBytes);
    m3JournalInfo info;
    if (need == 0 || snapshot == NULL || (journal == NULL && journalBytes > 0) || out == NULL ||
        capacity < need || !m3JournalDescribe(journalBytes > 0 ? journal : "", journalBytes, &info))
    {
        return 0;
    }
    uint8_t header[M3_REPLAY_HEADER];
    const uint32_t words[2] = {M3_REPLAY_MAGIC, M3_REPLAY_VERSION};
    const int32_t sizes[4] = {snapshotBytes, journalBytes, info.stepCount, info.opCount};
    memcpy(header, words, 8);
    memcpy(header + 8, sizes, 16);
    memcpy(header + 24, &finalHash, 8);
    uint8_t* p = (uint8_t*)out;
    memcpy(p, header, M3_REPLAY_HEADER);
    memcpy(p + M3_REPLAY_HEADER, snapshot, (size_t)snapshotBytes);
    if (journalBytes > 0)
    {
        memcpy(p + M3_REPLAY_HEADER + snapshotBytes, journal, (size_t)journalBytes);
    }
    return need;
}

bool m3ReplayDecode(const void* data, int32_t bytes, m3ReplayView* out)
{
    if (data == NULL || out == NULL || bytes < M3_REPLAY_HEADER)
    {
        return false;
    }
    const uint8_t* p = (const uint8_t*)data;
    uint32_t words[2];
    int32_t sizes[4];
    uint64_t finalHash;
    memcpy(words, p, 8);
    memcpy(sizes, p + 8, 16);
    memcpy(&finalHash, p + 24, 8);
    if (words[0] != M3_REPLAY_MAGIC || words[1] != M3_REPLAY_VERSION || sizes[0] <= 0 ||
        sizes[1] < 0 || (int64_t)M3_REPLAY_HEADER + sizes[0] + sizes[1] != (int64_t)bytes)
    {
        return false;
    }
    const uint8_t* journal = p + M3_REPLAY_HEADER + sizes[0];
    m3JournalInfo info;
    if (!m3JournalDescribe(journal, sizes[1], &info) || info.stepCount != sizes[2] ||
        info.opCount != sizes[3])
    {
        return false; // the header must describe its own stream
    }
    out->snapshot = p + M3_REPLAY_HEADER;
    out->snapshotBytes = sizes[0];
    out->journal = journal;
    out->journalBytes = sizes[1];
    out->finalHash = finalHash;
    out->info = info;
    return true;
}
